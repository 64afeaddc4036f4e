"""Differentiable scalar advection (mx_adv_scal_step / mx_adv_scal_step_boxes
and their hand-written adjoints), as torch.autograd.Functions.

* `ScalStepper(n, ncomp, dt, rel_eps, ppm_type=1, conservative=False)`: a
  C-ABI context for one fully periodic box [0, n)^3 (dx = 1/n, ppm_type 1 or
  2, MakeEdgeScal in the advective or the conservative form) with device
  buffers of the step; `step(s, u)` and `vjp(s, u, gbar)` take dense torch
  tensors:
  s (ncomp, n, n, n) indexed [c, k, j, i] and u = (ux, uy, uz), each
  (n, n, n) holding the faces 0..n-1 of its direction (face n is face 0).
* `BoxStepper(boxes, n, periodic, ncomp, dt, rel_eps, ppm_type=1,
  conservative=False)`: the same
  step over a box list that tiles the domain n = (nx, ny, nz), each
  direction periodic or outflow (foextrap); `step(s, u)` and
  `vjp(s, u, gbar)` take the same dense tensors as ScalStepper's, u[d]
  holding faces 0..n[d] along an outflow d; `step_mf` and `vjp_mf` run on
  the box MultiFabs themselves.
* `ScalStepFn.apply(stepper, s, ux, uy, uz)`: one step, differentiable in s
  and the velocity.
* `Rollout.apply(stepper, nsteps, every, s0, ux, uy, uz)`: nsteps steps with
  a fixed velocity, storing only every `every`-th state; the backward pass
  recomputes each segment forward from its checkpoint and runs the adjoint
  steps in reverse, accumulating the velocity gradient over the steps.
"""
import ctypes as C

import numpy as np

from . import _abi
from .fabs import Geometry, make_params


def periodic_mac_velocity(n, seed=0, nmodes=16, umax=1.0):
    """A discretely divergence-free MAC velocity on the periodic box, the
    discrete curl of a random smooth vector potential on cell edges, scaled to
    max |u| = umax.  Returns (ux, uy, uz), each (n, n, n) [k, j, i]."""
    rng = np.random.default_rng(seed)
    x = np.arange(n) / n
    h = np.arange(n) / n + 0.5 / n
    A = []
    for comp in range(3):
        a = np.zeros((n, n, n))
        # component comp lives at the mid-point along comp, on nodes elsewhere
        cz = h if comp == 2 else x
        cy = h if comp == 1 else x
        cx = h if comp == 0 else x
        Z, Y, X = np.meshgrid(cz, cy, cx, indexing="ij")
        for _ in range(nmodes):
            k = rng.integers(-3, 4, size=3)
            if not k.any():
                k[0] = 1
            a += rng.normal() * np.cos(2 * np.pi * (k[0] * X + k[1] * Y + k[2] * Z)
                                       + rng.uniform(0, 2 * np.pi))
        A.append(a)
    Ax, Ay, Az = A
    dx = 1.0 / n
    r = lambda a, ax: np.roll(a, -1, axis=ax)  # value at index + 1
    ux = (r(Az, 1) - Az) / dx - (r(Ay, 0) - Ay) / dx
    uy = (r(Ax, 0) - Ax) / dx - (r(Az, 2) - Az) / dx
    uz = (r(Ay, 2) - Ay) / dx - (r(Ax, 1) - Ax) / dx
    s = umax / max(np.abs(ux).max(), np.abs(uy).max(), np.abs(uz).max())
    return ux * s, uy * s, uz * s


def periodic_divergence(u):
    ux, uy, uz = u
    n = ux.shape[0]
    return (np.roll(ux, -1, 2) - ux + np.roll(uy, -1, 1) - uy + np.roll(uz, -1, 0) - uz) * n


class ScalStepper:
    """Device buffers and a C-ABI context for the step on [0, n)^3."""

    NG = {1: 3, 2: 4}  # ghost cells of the state the edge stencils read, per ppm_type

    def __init__(self, n, ncomp, dt, rel_eps, device="cuda:0", ppm_type=1, conservative=False):
        import torch
        import maestroex_amd as mx
        self.torch = torch
        self.n, self.ncomp, self.dt, self.rel_eps = n, ncomp, dt, rel_eps
        self.ppm_type = ppm_type
        self.cons = 1 if conservative else 0
        g = Geometry(3, [n] * 3, [1.0 / n] * 3, [1, 1, 1], [0, 0, 0], [0, 0, 0])
        p = make_params(g, ppm_type=ppm_type, nspec=ncomp, dt=dt, rel_eps=rel_eps)
        self.m = mx.gpu_maestro(p, [[((0, 0, 0), (n - 1, n - 1, n - 1))]])
        self.lib, self.fns = mx.load()
        dev = torch.device(device)
        if ppm_type not in self.NG:
            raise ValueError(f"ppm_type {ppm_type}: the step supports 1 and 2")
        self.ng = ng = self.NG[ppm_type]
        e = n + 2 * ng
        f64 = torch.float64
        self.s = torch.zeros((ncomp, e, e, e), dtype=f64, device=dev)
        self.snew = torch.zeros((ncomp, n, n, n), dtype=f64, device=dev)  # shape template
        self.u = []
        for d in range(3):
            shp = [n + 2, n + 2, n + 2]
            shp[2 - d] = n + 3  # nodal in d, one ghost face each side
            self.u.append(torch.zeros(shp, dtype=f64, device=dev))
        self._fab_s = self._fab(self.s, [-ng] * 3, ncomp)
        self._fab_u = (_abi.Fab * 3)(*[self._fab(self.u[d], [-1] * 3, 1) for d in range(3)])

    @staticmethod
    def _fab(t, begin, nc):
        f = _abi.Fab()
        nz, ny, nx = t.shape[-3:]
        f.p = t.data_ptr()
        f.jstride = nx
        f.kstride = nx * ny
        f.nstride = nx * ny * nz
        for d in range(3):
            f.begin[d] = begin[d]
        f.end[0], f.end[1], f.end[2] = begin[0] + nx, begin[1] + ny, begin[2] + nz
        f.ncomp = nc
        return f

    def _load(self, s, u):
        n, ng = self.n, self.ng
        self.s[:, ng:ng + n, ng:ng + n, ng:ng + n].copy_(s)
        for d in range(3):
            sl = [slice(1, 1 + n)] * 3
            self.u[d][tuple(sl)].copy_(u[d])

    def _check(self, rc, where):
        if rc != _abi.MX_OK:
            raise _abi.MxError(rc, where)

    def step(self, s, u):
        """snew for the state s (ncomp, n, n, n) and velocity u; the result is
        a new tensor, ordered on the library's stream (torch's default)."""
        self._load(s, u)
        snew = self.torch.empty_like(self.snew)
        self._check(self.fns["scal_step"](self.m.ctx, C.byref(self._fab_s), self._fab_u,
                                          C.byref(self._fab(snew, [0] * 3, self.ncomp)),
                                          self.ncomp, self.cons), "scal_step")
        return snew

    def vjp(self, s, u, gbar):
        """(sbar, [ubar_x, ubar_y, ubar_z]) for the output adjoint gbar: the
        adjoint reads the dense inputs in place (it needs no ghost cells) and
        writes new tensors."""
        s = s.contiguous()
        gbar = gbar.contiguous()
        u = [x.contiguous() for x in u]
        sb = self.torch.empty_like(s)
        ub = [self.torch.empty_like(x) for x in u]
        nc = self.ncomp
        fu = (_abi.Fab * 3)(*[self._fab(u[d], [0] * 3, 1) for d in range(3)])
        fub = (_abi.Fab * 3)(*[self._fab(ub[d], [0] * 3, 1) for d in range(3)])
        self._check(self.fns["scal_step_vjp"](self.m.ctx, C.byref(self._fab(s, [0] * 3, nc)), fu,
                                              C.byref(self._fab(gbar, [0] * 3, nc)),
                                              C.byref(self._fab(sb, [0] * 3, nc)), fub, nc,
                                              self.cons), "scal_step_vjp")
        return sb, ub


class BoxStepper:
    """Device MultiFabs and a C-ABI context for the step over a box list."""

    NG = {1: 3, 2: 4}

    def __init__(self, boxes, n, periodic, ncomp, dt, rel_eps, ppm_type=1, device="cuda:0",
                 spherical=0, conservative=False):
        import torch
        import maestroex_amd as mx
        from .fabs import Backend, FaceMultiFab, MultiFab
        if ppm_type not in self.NG:
            raise ValueError(f"ppm_type {ppm_type}: the step supports 1 and 2")
        self.torch = torch
        self.boxes = [tuple(map(tuple, b)) for b in boxes]
        self.n = list(n)
        self.periodic = [1 if p else 0 for p in periodic]
        self.ncomp, self.dt, self.rel_eps, self.ppm_type = ncomp, dt, rel_eps, ppm_type
        self.cons = 1 if conservative else 0
        lo = [_abi.PHYS_INTERIOR if p else _abi.PHYS_OUTFLOW for p in self.periodic]
        g = Geometry(3, self.n, [1.0 / q for q in self.n], self.periodic, lo, lo)
        p = make_params(g, ppm_type=ppm_type, nspec=ncomp, dt=dt, rel_eps=rel_eps,
                        species_pred_type=2, spherical=spherical)
        self.m = mx.gpu_maestro(p, [self.boxes])
        self.lib, self.fns = mx.load()
        self.bk = bk = Backend("torch", torch.device(device))
        ng = self.NG[ppm_type]
        self.s = MultiFab(self.boxes, ncomp, ng, 3, bk)
        self.snew = MultiFab(self.boxes, ncomp, 0, 3, bk)
        self.u = FaceMultiFab(self.boxes, 1, 1, 3, bk)
        self.g = MultiFab(self.boxes, ncomp, 0, 3, bk)
        self.sb = MultiFab(self.boxes, ncomp, 0, 3, bk)
        self.ub = FaceMultiFab(self.boxes, 1, 0, 3, bk)

    def _check(self, rc, where):
        if rc != _abi.MX_OK:
            raise _abi.MxError(rc, where)

    def step_mf(self, s, u, snew):
        """snew = step(s) on MultiFabs (s: >= 3 ghosts, 4 for ppm_type 2;
        u: >= 1 ghost face; both ghost layers are filled by the call)."""
        self._check(self.fns["scal_step_boxes"](self.m.ctx, C.byref(s.desc()), C.byref(u.desc()),
                                                C.byref(snew.desc()), self.ncomp, self.cons),
                    "scal_step_boxes")

    def vjp_mf(self, s, u, gbar, sbar, ubar):
        self._check(self.fns["scal_step_vjp_boxes"](self.m.ctx, C.byref(s.desc()),
                                                    C.byref(u.desc()), C.byref(gbar.desc()),
                                                    C.byref(sbar.desc()), C.byref(ubar.desc()),
                                                    self.ncomp, self.cons),
                    "scal_step_vjp_boxes")

    # dense global tensors <-> the box MultiFabs
    def _put_cells(self, mf, a):
        for b, (lo, hi) in enumerate(self.boxes):
            reg = mf.valid_region(b)
            mf.arrays[b][(slice(None),) + reg] = a[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1,
                                                   lo[0]:hi[0] + 1]

    def _get_cells(self, mf):
        n = self.n
        out = self.torch.empty((self.ncomp, n[2], n[1], n[0]), dtype=self.torch.float64,
                               device=self.bk.device)
        for b, (lo, hi) in enumerate(self.boxes):
            reg = mf.valid_region(b)
            out[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = \
                mf.arrays[b][(slice(None),) + reg]
        return out

    def _face_index(self, d, lo, hi):
        """Per direction (x, y, z) the global face indices of a box's faces
        lo..hi+1 along d (periodic: face n is face 0)."""
        idx = []
        for q in range(3):
            top = hi[q] + (1 if q == d else 0)
            r = self.torch.arange(lo[q], top + 1, device=self.bk.device)
            if q == d and self.periodic[q]:
                r = r % self.n[q]
            idx.append(r)
        return idx

    def _put_faces(self, fm, u):
        for d in range(3):
            for b, (lo, hi) in enumerate(self.boxes):
                ix, iy, iz = self._face_index(d, lo, hi)
                reg = fm[d].valid_region(b)
                fm[d].arrays[b][(0,) + reg] = u[d][iz][:, iy][:, :, ix]

    def _get_faces(self, fm):
        out = []
        for d in range(3):
            shp = [self.n[2], self.n[1], self.n[0]]
            if not self.periodic[d]:
                shp[2 - d] += 1
            a = self.torch.empty(shp, dtype=self.torch.float64, device=self.bk.device)
            for b, (lo, hi) in enumerate(self.boxes):
                ix, iy, iz = self._face_index(d, lo, hi)
                reg = fm[d].valid_region(b)
                v = fm[d].arrays[b][(0,) + reg]
                a[iz[:, None, None], iy[None, :, None], ix[None, None, :]] = v
            out.append(a)
        return out

    def step(self, s, u):
        self._put_cells(self.s, s)
        self._put_faces(self.u, u)
        self.step_mf(self.s, self.u, self.snew)
        return self._get_cells(self.snew)

    def vjp(self, s, u, gbar):
        self._put_cells(self.s, s)
        self._put_faces(self.u, u)
        self._put_cells(self.g, gbar)
        self.vjp_mf(self.s, self.u, self.g, self.sb, self.ub)
        return self._get_cells(self.sb), self._get_faces(self.ub)


def _autograd():
    import torch

    class ScalStepFn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, stepper, s, ux, uy, uz):
            ctx.stepper = stepper
            ctx.save_for_backward(s, ux, uy, uz)
            return stepper.step(s, (ux, uy, uz))

        @staticmethod
        def backward(ctx, g):
            s, ux, uy, uz = ctx.saved_tensors
            sb, ub = ctx.stepper.vjp(s, (ux, uy, uz), g.contiguous())
            return None, sb, ub[0], ub[1], ub[2]

    class Rollout(torch.autograd.Function):
        @staticmethod
        def forward(ctx, stepper, nsteps, every, s0, ux, uy, uz):
            u = (ux, uy, uz)
            ckpt = [s0.clone()]
            s = s0
            for k in range(nsteps):
                s = stepper.step(s, u)
                if (k + 1) % every == 0 and k + 1 < nsteps:
                    ckpt.append(s.clone())
            ctx.stepper, ctx.nsteps, ctx.every = stepper, nsteps, every
            ctx.ckpt = ckpt
            ctx.save_for_backward(ux, uy, uz)
            return s

        @staticmethod
        def backward(ctx, g):
            ux, uy, uz = ctx.saved_tensors
            u = (ux, uy, uz)
            st, nsteps, every = ctx.stepper, ctx.nsteps, ctx.every
            gb = g.contiguous()
            ubar = [torch.zeros_like(x) for x in u]
            nseg = (nsteps + every - 1) // every
            for seg in reversed(range(nseg)):
                k0 = seg * every
                k1 = min(k0 + every, nsteps)
                states = [ctx.ckpt[seg]]
                for _ in range(k0, k1 - 1):
                    states.append(st.step(states[-1], u))
                for s in reversed(states):
                    gb, ub = st.vjp(s, u, gb)
                    for d in range(3):
                        ubar[d] += ub[d]
            return None, None, None, gb, ubar[0], ubar[1], ubar[2]

    return ScalStepFn, Rollout


_FNS = {}


def __getattr__(name):
    # the autograd classes are built on first use, so that importing this
    # module does not import torch
    if name in ("ScalStepFn", "Rollout"):
        if not _FNS:
            _FNS["ScalStepFn"], _FNS["Rollout"] = _autograd()
        return _FNS[name]
    raise AttributeError(name)
