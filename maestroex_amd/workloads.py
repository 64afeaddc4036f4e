"""Seeded synthetic inputs for the advection pass.

The construction (one uint64 seed per workload):
  * domain [0,1]^D, n^D cells, uniform dx = 1/n, ghost width `ng`;
  * periodic in every direction but the vertical one (y in 2-D, z in 3-D),
    which is an outflow boundary (foextrap ghost cells), unless
    `all_periodic`;
  * rho = exp(-r/0.3) * (1 + 0.1 exp(-d^2/0.05^2)): a Gaussian bubble at the
    domain centre on a hydrostatic exponential background (multiplicative,
    following the exp(-d^2/W^2) form of the reference's test_advect
    initial data, Exec/unit_tests/test_advect/MaestroInitData.cpp:43-44);
  * rho0 is the horizontal average of rho; the advected density is
    rho' = rho - rho0 (PutInPertForm);
  * h = 2 - r (a smooth enthalpy), rhoh = rho h, rhoh0 its horizontal average,
    the advected enthalpy is (rho h)' = rhoh - rhoh0;
  * NumSpec mass fractions uniform random in [0,1], normalised to sum 1 (the
    advected species are X, after ConvertRhoXToX);
  * velocity = sum of `nmodes` Fourier modes c_m cos(2 pi k_m.x + phi_m) with
    c_m perpendicular to k_m (divergence free), integer wavenumbers in the
    periodic directions, scaled to max|u| = 1;
  * w0(r) = 0.01 * r_edge (planar base-state velocity at cell edges),
    w0_cart its cell-centred average (Put1dArrayOnCart), ufull = utilde +
    w0_cart;
  * smooth velocity and scalar forcing terms;
  * dt from CFL 0.5 on the full velocity, rel_eps = 1e-8 * umax.
Ghost cells are filled as FillPatch would: periodic images in periodic
directions, first-order extrapolation (copy of the nearest valid cell) at
the outflow boundary.
"""
import sys

import numpy as np

from . import _abi
from .fabs import BaseState, FaceMultiFab, Geometry, MultiFab, make_params, split_domain


def cell_to_edge(s0):
    """Maestro::CelltoEdge (MaestroCelltoEdge.cpp:5-55), one radial level."""
    nr = len(s0)
    e = np.zeros(nr + 1)
    for r in range(nr + 1):
        if r == 0:
            e[r] = s0[r]
        elif r == 1:
            e[r] = 0.5 * (s0[r - 1] + s0[r])
        elif r == nr:
            e[r] = s0[r - 1]
        elif r == nr - 1:
            e[r] = 0.5 * (s0[r] + s0[r - 1])
        else:
            tmp = 7.0 / 12.0 * (s0[r] + s0[r - 1]) - 1.0 / 12.0 * (s0[r + 1] + s0[r - 2])
            lo = min(s0[r], s0[r - 1])
            hi = max(s0[r], s0[r - 1])
            e[r] = min(max(tmp, lo), hi)
    return e


class Workload:
    """All inputs and outputs of one level of the pass, on one backend."""

    def __init__(self, dim=3, n=32, max_grid=16, ppm_type=1, nspec=3, seed=1234, nmodes=None,
                 ng=4, all_periodic=False, backend=None, ppm_trace_forces=0,
                 phys_bc=None, zero_velocity=False, shift=(0, 0, 0), spherical=False,
                 use_exact_base_state=0, species_pred_type=1, enthalpy_pred_type=1,
                 boxes=None):
        from .fabs import Backend
        self.backend = backend or Backend("numpy")
        self.dim = dim
        self.n = n
        self.ng = ng
        nm = nmodes if nmodes is not None else (8 if dim == 2 else 16)
        vert = dim - 1
        periodic = [1] * dim
        lo_bc = [_abi.PHYS_INTERIOR] * dim
        hi_bc = [_abi.PHYS_INTERIOR] * dim
        if spherical:
            # a star in a box: outflow on every side (3-D only)
            assert dim == 3
            periodic = [0] * dim
            lo_bc = [_abi.PHYS_OUTFLOW] * dim
            hi_bc = [_abi.PHYS_OUTFLOW] * dim
        elif not all_periodic:
            periodic[vert] = 0
            lo_bc[vert] = _abi.PHYS_OUTFLOW
            hi_bc[vert] = _abi.PHYS_OUTFLOW
        if phys_bc is not None:
            # explicit physical boundaries: {dir: (lo, hi)}; non-interior => not periodic
            for d, (bl, bh) in phys_bc.items():
                lo_bc[d], hi_bc[d] = bl, bh
                periodic[d] = 1 if bl == _abi.PHYS_INTERIOR else 0
        self.geom = Geometry(dim, [n] * dim, [1.0 / n] * dim, periodic, lo_bc, hi_bc)
        # boxes: an explicit box list [((lo), (hi)), ...] covering the domain,
        # else boxes of max_grid cells
        self.boxes = list(boxes) if boxes is not None else split_domain(self.geom, max_grid)
        rng = np.random.default_rng(seed)

        nscal = 4 + nspec
        N = n
        dx = 1.0 / n
        ext = N + 2 * ng
        # global index arrays including ghosts, mapped to valid cells
        gidx = np.arange(-ng, N + ng)

        shift = list(shift) + [0] * (3 - len(shift))

        def mapped(d, sh):
            g = gidx - sh
            if periodic[d]:
                return np.mod(g, N)
            return np.clip(g, 0, N - 1)

        mi = [mapped(d, shift[d]) for d in range(dim)]
        # the horizontal averages come from the unshifted field, so that a
        # periodic shift of the input leaves the base state bit-identical
        mi0 = [mapped(d, 0) for d in range(dim)]
        xc = [(mi[d] + 0.5) * dx for d in range(dim)]

        def grid(arrs):
            # broadcast per-direction 1-D arrays to (z, y, x) or (y, x)
            if dim == 3:
                return arrs[0][None, None, :], arrs[1][None, :, None], arrs[2][:, None, None]
            return arrs[0][None, :], arrs[1][:, None]

        def radius(Xg):
            if spherical:
                return np.sqrt(sum((Xg[d] - 0.5) ** 2 for d in range(dim)))
            return Xg[vert]

        def density(Xg):
            rr = radius(Xg)
            dist2 = sum((Xg[d] - 0.5) ** 2 for d in range(dim))
            rho_ = np.exp(-rr / 0.3) * (1.0 + 0.1 * np.exp(-dist2 / 0.05 ** 2))
            rho_ = np.broadcast_to(rho_, (ext,) * dim).copy()
            return rho_, rho_ * np.broadcast_to(2.0 - rr, (ext,) * dim)

        X = grid(xc)
        rho, rhoh = density(X)
        valid = tuple(slice(ng, ng + N) for _ in range(dim))
        haxes = tuple(a for a in range(dim) if a != 0)  # all but vertical (axis 0 = vertical)
        rho_u, rhoh_u = density(grid([(mi0[d] + 0.5) * dx for d in range(dim)]))
        rho0 = rho_u[valid].mean(axis=haxes)
        rhoh0 = rhoh_u[valid].mean(axis=haxes)
        rmap = mi[vert]
        shp = [1] * dim
        shp[0] = ext
        rhop = rho - rho0[rmap].reshape(shp)
        rhohp = rhoh - rhoh0[rmap].reshape(shp)

        Xs = rng.random((nspec,) + (N,) * dim)
        Xs /= Xs.sum(axis=0, keepdims=True)
        ix = np.ix_(*[mi[d] for d in reversed(range(dim))])
        Xg = np.stack([Xs[k][ix] for k in range(nspec)])

        # velocity: divergence-free Fourier modes
        u = np.zeros((dim,) + (ext,) * dim)
        if not zero_velocity:
            for m in range(nm):
                k = np.zeros(dim)
                for d in range(dim):
                    if periodic[d]:
                        k[d] = rng.integers(-3, 4)
                    else:
                        k[d] = rng.uniform(-2.0, 2.0)
                if not np.any(k):
                    k[0] = 1.0
                a = rng.normal(size=dim)
                if dim == 2:
                    cvec = np.array([-k[1], k[0]]) * rng.normal()
                else:
                    cvec = np.cross(k, a)
                phi = rng.uniform(0, 2 * np.pi)
                ph = [np.exp(1j * 2 * np.pi * k[d] * xc[d]) for d in range(dim)]
                G = grid(ph)
                mode = G[0] * G[1] * (G[2] if dim == 3 else 1.0) * np.exp(1j * phi)
                cm = np.real(mode)
                for d in range(dim):
                    u[d] += cvec[d] * cm
            umax0 = np.max(np.abs(u[(slice(None),) + valid]))
            u /= umax0

        nr = N
        r_edge = np.arange(nr + 1) * dx
        # a planar w0 is a function of height; with a periodic vertical
        # direction it must vanish (the faces r = 0 and r = nr coincide)
        w0 = np.zeros(nr + 1) if (periodic[vert] or spherical) else 0.01 * r_edge
        w0c = 0.5 * (w0[:-1] + w0[1:])
        w0_cart = np.zeros_like(u)
        if spherical:
            # radial w0(r) = 0.05 r (1 - r) along the unit radial vector
            rc = radius(X)
            wr = 0.05 * rc * (1.0 - rc)
            with np.errstate(invalid="ignore", divide="ignore"):
                for d in range(dim):
                    w0_cart[d] = np.broadcast_to(np.where(rc > 0, wr * (X[d] - 0.5) / rc, 0.0),
                                                 (ext,) * dim)
        else:
            w0_cart[vert] = np.broadcast_to(w0c[rmap].reshape(shp), (ext,) * dim)
        ufull = u + w0_cart

        umax = float(np.max(np.abs(ufull[(slice(None),) + valid])))
        if umax == 0.0:
            umax = 1.0
        self.dt = 0.5 * dx / umax
        self.rel_eps = 1.0e-8 * umax

        # forcing
        arg = 2 * np.pi * sum((d + 1) * X[d] for d in range(dim))
        vf = np.stack([0.05 * np.cos(arg + c) * np.ones((ext,) * dim) for c in range(dim)])
        vf[vert] += -0.1 * rhop / rho
        sf = np.stack([0.02 * np.sin(arg + 0.5 * c) * np.ones((ext,) * dim) for c in range(nscal)])

        scal = np.zeros((nscal,) + (ext,) * dim)
        scal[0] = rhop
        scal[1] = rhohp
        scal[4:4 + nspec] = Xg

        self.params = make_params(self.geom, ppm_type=ppm_type, nspec=nspec, dt=self.dt,
                                  rel_eps=self.rel_eps, ppm_trace_forces=ppm_trace_forces,
                                  spherical=1 if spherical else 0,
                                  use_exact_base_state=use_exact_base_state,
                                  species_pred_type=species_pred_type,
                                  enthalpy_pred_type=enthalpy_pred_type)
        bk = self.backend
        fng = ng if ppm_trace_forces else 1

        def cellmf(data, ncomp, ngrow):
            mf = MultiFab(self.boxes, ncomp, ngrow, dim, bk)
            for b, (lo, hi) in enumerate(self.boxes):
                sl = [slice(lo[d] - ngrow + ng, hi[d] + ngrow + 1 + ng) for d in reversed(range(dim))]
                sub = data[(slice(None),) + tuple(sl)]
                if dim == 2:
                    sub = sub[:, None, :, :]
                mf.arrays[b] = bk.from_numpy(sub)
            return mf

        self.utilde = cellmf(u, dim, ng)
        self.ufull = cellmf(ufull, dim, ng)
        self.w0_cart = cellmf(w0_cart, dim, min(2, ng))
        self.vel_force = cellmf(vf, dim, fng)
        self.scal = cellmf(scal, nscal, ng)
        self.scal_force = cellmf(sf, nscal, fng)

        self.w0 = BaseState(w0, bk)
        self.rho0 = BaseState(rho0, bk)
        self.rhoh0 = BaseState(rhoh0, bk)
        self.rho0_edge = BaseState(cell_to_edge(rho0), bk)
        self.rhoh0_edge = BaseState(cell_to_edge(rhoh0), bk)
        self.rho0_predicted_edge = BaseState(cell_to_edge(rho0), bk)

        # spherical: w0 and the base states on faces (MakeW0mac / MakeS0mac
        # style: the radial profiles evaluated at each face centre)
        self.w0mac = self.r0mac = self.rh0mac = self.h0mac = None
        if spherical:
            def face_field(fn):
                fm = FaceMultiFab(self.boxes, 1, 1, dim, bk)
                for d in range(dim):
                    for b in range(len(self.boxes)):
                        def g(ii, jj, kk, d=d):
                            idx = [ii, jj, kk]
                            xs = [(idx[q] + (0.0 if q == d else 0.5)) * dx for q in range(3)]
                            Z, Y, Xx = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
                            return fn(d, (Xx, Y, Z))[None]
                        fm[d].set_from_global(b, g)
                return fm

            def rad(P):
                return np.sqrt(sum((P[q] - 0.5) ** 2 for q in range(3)))

            def w0f(d, P):
                r = rad(P)
                with np.errstate(invalid="ignore", divide="ignore"):
                    return np.where(r > 0, 0.05 * r * (1.0 - r) * (P[d] - 0.5) / r, 0.0)

            self.w0mac = face_field(w0f)
            self.r0mac = face_field(lambda d, P: np.exp(-rad(P) / 0.3))
            self.h0mac = face_field(lambda d, P: 2.0 - rad(P))
            self.rh0mac = face_field(lambda d, P: np.exp(-rad(P) / 0.3) * (2.0 - rad(P)))

        self.nscal = nscal
        self.alloc_outputs()

    def alloc_outputs(self, fill=0.0):
        bk = self.backend
        dim = self.dim
        self.utrans = FaceMultiFab(self.boxes, 1, 1, dim, bk, fill=fill)
        self.umac = FaceMultiFab(self.boxes, 1, 1, dim, bk, fill=fill)
        self.sedge = FaceMultiFab(self.boxes, self.nscal, 0, dim, bk, fill=fill)
        self.sflux = FaceMultiFab(self.boxes, self.nscal, 0, dim, bk, fill=fill)
        self.etarhoflux = MultiFab(self.boxes, 1, 1, dim, bk, nodal_dir=dim - 1, fill=fill)

    def pass_fields(self):
        return dict(utilde=self.utilde, ufull=self.ufull, vel_force=self.vel_force,
                    w0_cart=self.w0_cart, scal=self.scal, scal_force=self.scal_force,
                    w0mac=self.w0mac, r0mac=self.r0mac, rh0mac=self.rh0mac, h0mac=self.h0mac,
                    w0=self.w0, rho0=self.rho0, rho0_edge=self.rho0_edge,
                    rhoh0=self.rhoh0, rhoh0_edge=self.rhoh0_edge,
                    rho0_predicted_edge=self.rho0_predicted_edge, utrans=self.utrans,
                    umac=self.umac, sedge=self.sedge, sflux=self.sflux,
                    etarhoflux=self.etarhoflux)

    def valid_cells(self):
        return self.n ** self.dim


# ---------------------------------------------------------------------------
# An uneven AMR hierarchy (every level a list of boxes, fields per box)

def _hash_uniform(seed, lev, comp, i, j, k):
    """Deterministic uniform [0, 1) per (seed, level, component, cell): a
    splitmix64 finaliser over the packed integer coordinates."""
    with np.errstate(over="ignore"):
        x = (np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15)
             + np.uint64(lev) * np.uint64(0xBF58476D1CE4E5B9)
             + np.uint64(comp) * np.uint64(0x94D049BB133111EB))
        x = (x + (i.astype(np.uint64) << np.uint64(42)) + (j.astype(np.uint64) << np.uint64(21))
             + k.astype(np.uint64))
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    return (x >> np.uint64(11)).astype(np.float64) / float(1 << 53)


def uneven_boxes(n, parts, lo, hi, seed=0):
    """The domain [0, n)^3 as parts^3 boxes whose edges, per direction, are a
    seeded random split of n into `parts` lengths in [lo, hi]."""
    rng = np.random.default_rng(seed)
    cuts = []
    for _ in range(3):
        lens = np.full(parts, n // parts)
        lens[: n - lens.sum()] += 1
        for _ in range(8 * parts):
            a, b = rng.integers(0, parts, size=2)
            m = int(rng.integers(1, 9))
            if a != b and lens[a] - m >= lo and lens[b] + m <= hi:
                lens[a] -= m
                lens[b] += m
        edges = np.concatenate([[0], np.cumsum(lens)])
        cuts.append([(int(edges[q]), int(edges[q + 1]) - 1) for q in range(parts)])
    return [((x0, y0, z0), (x1, y1, z1)) for (z0, z1) in cuts[2] for (y0, y1) in cuts[1]
            for (x0, x1) in cuts[0]]


def _partition(rng, n, lo, hi):
    """Split [0, n) into consecutive segments of length lo..hi."""
    segs = []
    start = 0
    while start < n:
        rest = n - start
        if rest <= hi:
            if rest >= lo or not segs:
                segs.append((start, n - 1))
                break
            # merge a short remainder into the previous segment
            a, _ = segs.pop()
            segs.append((a, n - 1))
            break
        L = int(rng.integers(lo, hi + 1))
        if rest - L < lo:
            L = rest - lo
        segs.append((start, start + L - 1))
        start += L
    return segs


def _random_boxes(rng, n, count, lo, hi, region_frac=1.0, align=1, tries=200000):
    """`count` non-overlapping boxes with random edges lo..hi inside the
    first region_frac of [0, n)^3 along every axis; corners and edges are
    multiples of `align` (2: every box coarsens by the refinement ratio onto
    whole coarse cells, as AMReX's blocking factor guarantees)."""
    extent = max(hi, int(round(n * region_frac)))
    occ = []
    boxes = []
    t = 0
    while len(boxes) < count and t < tries:
        t += 1
        e = [align * int(rng.integers(-(-lo // align), hi // align + 1)) for _ in range(3)]
        a = [align * int(rng.integers(0, (extent - e[d]) // align + 1)) for d in range(3)]
        b = [a[d] + e[d] - 1 for d in range(3)]
        if any(all(a[d] <= q[1][d] and q[0][d] <= b[d] for d in range(3)) for q in occ):
            continue
        occ.append((a, b))
        boxes.append((tuple(a), tuple(b)))
    return boxes


def _nested_boxes(rng, parents, count, lo, hi, ratio=2, align=2, buf=2, tries=200000):
    """`count` non-overlapping boxes with random edges lo..hi (multiples of
    `align`, at even corners), each inside one refined parent box shrunk by
    `buf` fine cells."""
    occ = []
    boxes = []
    t = 0
    while len(boxes) < count and t < tries:
        t += 1
        plo, phi = parents[int(rng.integers(0, len(parents)))]
        room = [ratio * (phi[d] - plo[d] + 1) - 2 * buf for d in range(3)]
        top = [min(hi, room[d]) // align for d in range(3)]
        if any(top[d] * align < lo for d in range(3)):
            continue
        e = [align * int(rng.integers(-(-lo // align), top[d] + 1)) for d in range(3)]
        a = [ratio * plo[d] + buf + align * int(rng.integers(0, (room[d] - e[d]) // align + 1))
             for d in range(3)]
        b = [a[d] + e[d] - 1 for d in range(3)]
        if any(all(a[d] <= q[1][d] and q[0][d] <= b[d] for d in range(3)) for q in occ):
            continue
        occ.append((a, b))
        boxes.append((tuple(a), tuple(b)))
    return boxes


class Hierarchy:
    """An uneven AMR hierarchy with the field construction of Workload,
    evaluated per box (no full-level arrays), refinement ratio 2.

    level 0: n0^3 split into boxes whose edges are lo0..hi0 along each axis;
    level 1: `counts[0]` non-overlapping boxes with random even edges
    `edges[0]` at even corners inside the first `fracs[0]` of the level's
    index space (each coarsens onto whole cells of level 0);
    level l > 1: `counts[l-1]` non-overlapping boxes with random even edges
    `edges[l-1]` at even corners, each inside one box of level l - 1 with at
    least one coarse cell of that box around it (proper nesting: every ghost
    face of the level lies over a coarser-level face).
    Periodic in x and y, outflow in z; w0 = 0.01 z on every level; one dt
    (CFL 0.5 on the finest level) for all levels.  Every level follows the
    single-level semantics of the pass: ghost cells hold the analytic fields
    (periodic images / nearest valid cell in z), as FillPatch from a fine
    enough level would provide.
    """

    def __init__(self, n0=32, level0_edges=(8, 16), counts=(6, 8), edges=((8, 16), (4, 12)),
                 fracs=(0.6, 0.4), ppm_type=1, nspec=3, seed=7, nmodes=16, ng=4, backend=None,
                 verbose=False, reflux_type=1):
        from .fabs import Backend
        self.verbose = verbose
        self.backend = bk = backend or Backend("numpy")
        self.dim = 3
        rng = np.random.default_rng(seed)
        nlev = 1 + len(counts)
        self.nlev = nlev
        self.ns = [n0 * (2 ** l) for l in range(nlev)]
        segs = [_partition(rng, n0, *level0_edges) for _ in range(3)]
        b0 = [((x[0], y[0], z[0]), (x[1], y[1], z[1])) for z in segs[2] for y in segs[1]
              for x in segs[0]]
        self.boxes = [b0]
        for l in range(1, nlev):
            if l == 1:
                self.boxes.append(_random_boxes(rng, self.ns[l], counts[l - 1], *edges[l - 1],
                                                region_frac=fracs[l - 1], align=2))
            else:
                self.boxes.append(_nested_boxes(rng, self.boxes[l - 1], counts[l - 1],
                                                *edges[l - 1]))
        geoms = [Geometry(3, [n] * 3, [1.0 / n] * 3, [1, 1, 0],
                          [_abi.PHYS_INTERIOR, _abi.PHYS_INTERIOR, _abi.PHYS_OUTFLOW],
                          [_abi.PHYS_INTERIOR, _abi.PHYS_INTERIOR, _abi.PHYS_OUTFLOW])
                 for n in self.ns]

        # divergence-free Fourier modes, integer wavenumbers in x, y
        self.modes = []
        for _ in range(nmodes):
            k = np.array([rng.integers(-3, 4), rng.integers(-3, 4), rng.uniform(-2.0, 2.0)],
                         dtype=float)
            if not np.any(k):
                k[0] = 1.0
            cvec = np.cross(k, rng.normal(size=3))
            self.modes.append((k, cvec, rng.uniform(0, 2 * np.pi)))
        # scale so that max |u| over the finest level's cell centres is 1
        # (estimated on a coarse sample of the domain)
        xs = (np.arange(64) + 0.5) / 64
        X = np.meshgrid(xs, xs, xs, indexing="ij")
        self.uscale = 1.0
        self.uscale = 1.0 / np.max(np.abs(self._velocity(X)))
        umax = 1.0 + 0.01
        dxf = 1.0 / self.ns[-1]
        self.dt = 0.5 * dxf / umax
        self.rel_eps = 1.0e-8 * umax
        self.nscal = 4 + nspec
        self.nspec = nspec
        self.ng = ng
        self.seed = seed
        self.params = make_params(geoms[0], ppm_type=ppm_type, nspec=nspec, dt=self.dt,
                                  rel_eps=self.rel_eps, extra_levels=geoms[1:],
                                  reflux_type=reflux_type)
        self.levels = [self._level(l) for l in range(nlev)]

    # analytic fields -----------------------------------------------------
    def _velocity(self, X):
        u = np.zeros((3,) + np.broadcast(*X).shape)
        for k, cvec, phi in self.modes:
            arg = 2 * np.pi * (k[0] * X[0] + k[1] * X[1] + k[2] * X[2]) + phi
            c = np.cos(arg)
            for d in range(3):
                u[d] += cvec[d] * c
        return u * self.uscale

    def _velocity_cached(self, X, key):
        # utilde and ufull of one box are evaluated on the same cells
        if getattr(self, "_vkey", None) != key:
            self._vkey = key
            self._vval = self._velocity(X)
        return self._vval

    def _level(self, l):
        n = self.ns[l]
        dx = 1.0 / n
        bk = self.backend
        boxes = self.boxes[l]
        ng = self.ng
        nscal = self.nscal
        out = {}

        def cellmf(fn, ncomp, ngrow):
            mf = MultiFab(boxes, ncomp, ngrow, 3, bk)
            for b, (lo, hi) in enumerate(boxes):
                idx = [np.arange(lo[d] - ngrow, hi[d] + ngrow + 1) for d in range(3)]
                # ghost cells: periodic images in x, y; nearest valid cell in z
                mi = [np.mod(idx[0], n), np.mod(idx[1], n), np.clip(idx[2], 0, n - 1)]
                K, J, I = np.meshgrid(mi[2], mi[1], mi[0], indexing="ij")
                Xc = ((I + 0.5) * dx, (J + 0.5) * dx, (K + 0.5) * dx)
                self._box_key = (l, b, ngrow)
                mf.arrays[b] = bk.from_numpy(np.ascontiguousarray(fn(Xc, I, J, K)))
            return mf

        def rho_and_rhoh(Xc):
            z = Xc[2]
            dist2 = sum((Xc[d] - 0.5) ** 2 for d in range(3))
            rho = np.exp(-z / 0.3) * (1.0 + 0.1 * np.exp(-dist2 / 0.05 ** 2))
            return rho, rho * (2.0 - z)

        def scal(Xc, I, J, K):
            rho, rhoh = rho_and_rhoh(Xc)
            z = Xc[2]
            rho0 = np.exp(-z / 0.3)
            s = np.zeros((nscal,) + rho.shape)
            s[0] = rho - rho0
            s[1] = rhoh - rho0 * (2.0 - z)
            xs = np.stack([_hash_uniform(self.seed, l, c, I, J, K) for c in range(self.nspec)])
            s[4:] = xs / xs.sum(axis=0, keepdims=True)
            return s

        def w0c(Xc):
            w = np.zeros((3,) + np.broadcast(*Xc).shape)
            w[2] = 0.01 * Xc[2]
            return w

        def vel(Xc, I, J, K):
            return self._velocity_cached(Xc, self._box_key)

        def ufull(Xc, I, J, K):
            return self._velocity_cached(Xc, self._box_key) + w0c(Xc)

        def vforce(Xc, I, J, K):
            arg = 2 * np.pi * (Xc[0] + 2 * Xc[1] + 3 * Xc[2])
            f = np.stack([0.05 * np.cos(arg + c) for c in range(3)])
            rho, _ = rho_and_rhoh(Xc)
            f[2] += -0.1 * (rho - np.exp(-Xc[2] / 0.3)) / rho
            return f

        def sforce(Xc, I, J, K):
            arg = 2 * np.pi * (Xc[0] + 2 * Xc[1] + 3 * Xc[2])
            return np.stack([0.02 * np.sin(arg + 0.5 * c) for c in range(nscal)])

        fng = 1
        if self.verbose:
            print(f"[hierarchy] level {l}: {len(boxes)} boxes", file=sys.stderr, flush=True)
        out["utilde"] = cellmf(vel, 3, ng)
        out["ufull"] = cellmf(ufull, 3, ng)
        self._vkey = None
        out["w0_cart"] = cellmf(lambda Xc, I, J, K: w0c(Xc), 3, min(2, ng))
        out["vel_force"] = cellmf(vforce, 3, fng)
        out["scal"] = cellmf(scal, nscal, ng)
        out["scal_force"] = cellmf(sforce, nscal, fng)
        zc = (np.arange(n) + 0.5) * dx
        ze = np.arange(n + 1) * dx
        nrow = self.ns[-1] + 1  # one row length for every level's base states
        rows = lambda v: np.pad(v, (0, nrow - len(v)))[None, :]
        base = np.zeros((self.nlev, nrow))

        def bs(v):
            a = base.copy()
            a[l] = rows(v)[0]
            return BaseState(a, bk)

        rho0 = np.exp(-zc / 0.3)
        rhoh0 = rho0 * (2.0 - zc)
        out["w0"] = bs(0.01 * ze)
        out["rho0"] = bs(rho0)
        out["rhoh0"] = bs(rhoh0)
        out["rho0_edge"] = bs(cell_to_edge(rho0))
        out["rhoh0_edge"] = bs(cell_to_edge(rhoh0))
        out["rho0_predicted_edge"] = bs(cell_to_edge(rho0))
        out["w0mac"] = out["r0mac"] = out["rh0mac"] = out["h0mac"] = None
        out["utrans"] = FaceMultiFab(boxes, 1, 1, 3, bk)
        out["umac"] = FaceMultiFab(boxes, 1, 1, 3, bk)
        out["sedge"] = FaceMultiFab(boxes, nscal, 0, 3, bk)
        out["sflux"] = FaceMultiFab(boxes, nscal, 0, 3, bk)
        out["etarhoflux"] = MultiFab(boxes, 1, 1, 3, bk, nodal_dir=2)
        return out

    def pass_fields(self, l):
        return self.levels[l]

    def valid_cells(self):
        return sum(int(np.prod([hi[d] - lo[d] + 1 for d in range(3)]))
                   for bl in self.boxes for (lo, hi) in bl)
