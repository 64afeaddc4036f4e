"""Python mirror of the Maestro member functions on the advection path.

`Maestro` holds a C-ABI context (the Maestro members the path reads) and
exposes the reference's function names with the reference's argument
meaning.  It marshals MultiFab containers (maestroex_amd.fabs) into the
C-ABI descriptors and raises MxError on a non-zero return code, where the
reference would call amrex::Abort / Error.
"""
import ctypes as C

import numpy as np

from . import _abi
from .fabs import boxes_desc, empty_face_desc, empty_mf_desc


def _face(f):
    return f.desc() if f is not None else empty_face_desc()


def _mf(m):
    return m.desc() if m is not None else empty_mf_desc()


def _bs(b):
    return b.desc() if b is not None else _abi.BaseStateDesc()


class Maestro:
    def __init__(self, fns, params, levels, name="mx_adv", keepalive=None):
        self.fns = fns
        self.name = name
        self.params = params
        self._keep = keepalive
        self.ctx = C.c_void_p()
        self._check(fns["create"](C.byref(params), C.byref(self.ctx)), "create")
        self.levels = levels
        for lev, boxes in enumerate(levels):
            bd = boxes_desc(boxes)
            self._check(fns["set_level"](self.ctx, lev, len(boxes), bd), "set_level")

    def _check(self, rc, where):
        if rc != _abi.MX_OK:
            raise _abi.MxError(rc, f"{self.name}.{where}")

    def close(self):
        if self.ctx:
            self.fns["destroy"](self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- members ---------------------------------------------------------
    def set_dt(self, dt, rel_eps):
        self.params.dt = dt
        self.params.rel_eps = rel_eps
        self._check(self.fns["set_dt"](self.ctx, dt, rel_eps), "set_dt")

    def bc_setup(self):
        """Maestro::BCSetup -> (bcs_s, bcs_u) as ctypes BCRec arrays."""
        bs = (_abi.BCRec * self.params.nscal)()
        bu = (_abi.BCRec * 3)()
        self._check(self.fns["bc_setup"](self.ctx, bs, bu), "bc_setup")
        return bs, bu

    def set_tuning(self, key, value):
        """Kernel selection / launch shape; results are identical for every choice."""
        self._check(self.fns["set_tuning"](self.ctx, key.encode(), int(value)), "set_tuning")

    def sync(self):
        if "sync" in self.fns:
            self._check(self.fns["sync"](self.ctx), "sync")

    # ---- unit-level ------------------------------------------------------
    def PPM(self, lev, bx, s, u, v, w, Ip, Im, bcs, is_umac, comp, bccomp):
        """Maestro::PPM on one box; s,u,v,w,Ip,Im are (MultiFab, box) pairs."""
        b = _abi.Box()
        for d in range(3):
            b.lo[d], b.hi[d] = bx[0][d], bx[1][d]
        fs = [m.fab_desc(i) if m is not None else None for (m, i) in (s, u, v, w, Ip, Im)]
        ptrs = [C.byref(f) if f is not None else None for f in fs]
        self._check(self.fns["ppm"](self.ctx, lev, C.byref(b), *ptrs, bcs, len(bcs),
                                    int(is_umac), comp, bccomp), "PPM")

    def Slope(self, lev, d, bx, s, sl, bcs, ncomp, bc_start_comp, slope_order):
        """Maestro::Slopex/Slopey/Slopez on one box."""
        b = _abi.Box()
        for q in range(3):
            b.lo[q], b.hi[q] = bx[0][q], bx[1][q]
        fs = s[0].fab_desc(s[1])
        fl = sl[0].fab_desc(sl[1])
        self._check(self.fns["slope"](self.ctx, lev, d, C.byref(b), C.byref(fs), C.byref(fl),
                                      bcs, len(bcs), ncomp, bc_start_comp, slope_order), "Slope")

    # ---- level-wide stages -----------------------------------------------
    def MakeUtrans(self, lev, utilde, ufull, utrans, w0mac, w0_cart):
        a = [_mf(utilde), _mf(ufull), _face(utrans), _face(w0mac), _mf(w0_cart)]
        self._check(self.fns["make_utrans"](self.ctx, lev, *[C.byref(x) for x in a]), "MakeUtrans")

    def Addw0(self, lev, uedge, w0mac, w0, mult):
        a = [_face(uedge), _face(w0mac), _bs(w0)]
        self._check(self.fns["addw0"](self.ctx, lev, *[C.byref(x) for x in a], float(mult)), "Addw0")

    def FillFaceGhosts(self, lev, uedge):
        u = _face(uedge)
        self._check(self.fns["fill_face_ghosts"](self.ctx, lev, C.byref(u)), "FillFaceGhosts")

    def VelPred(self, lev, utilde, ufull, utrans, umac, w0mac, force, w0_cart):
        a = [_mf(utilde), _mf(ufull), _face(utrans), _face(umac), _face(w0mac), _mf(force),
             _mf(w0_cart)]
        self._check(self.fns["velpred"](self.ctx, lev, *[C.byref(x) for x in a]), "VelPred")

    def MakeEdgeScal(self, lev, state, sedge, umac, force, is_vel, bcs, start_scomp,
                     start_bccomp, num_comp, is_conservative):
        a = [_mf(state), _face(sedge), _face(umac), _mf(force)]
        self._check(self.fns["make_edge_scal"](self.ctx, lev, *[C.byref(x) for x in a],
                                               int(is_vel), bcs, len(bcs), start_scomp,
                                               start_bccomp, num_comp, int(is_conservative)),
                    "MakeEdgeScal")

    # ---- base state <-> Cartesian ---------------------------------------
    def Put1dArrayOnCart(self, lev, s0, s0_cart, is_input_edge_centered, is_output_a_vector,
                         sg=None):
        """Maestro::Put1dArrayOnCart on one level (sg: SphGeom, spherical only)."""
        a = [_bs(s0), _mf(s0_cart)]
        self._check(self.fns["put_1d_array_on_cart"](
            self.ctx, lev, *[C.byref(x) for x in a], int(is_input_edge_centered),
            int(is_output_a_vector), C.byref(sg) if sg is not None else None), "Put1dArrayOnCart")

    def MakeW0mac(self, lev, w0, w0_cart, w0mac, sg):
        a = [_bs(w0), _mf(w0_cart), _face(w0mac)]
        self._check(self.fns["make_w0mac"](self.ctx, lev, *[C.byref(x) for x in a],
                                           C.byref(sg)), "MakeW0mac")

    def MakeS0mac(self, lev, s0, s0_cart, s0mac, sg):
        a = [_bs(s0), _mf(s0_cart), _face(s0mac)]
        self._check(self.fns["make_s0mac"](self.ctx, lev, *[C.byref(x) for x in a],
                                           C.byref(sg)), "MakeS0mac")

    def PutInPertForm(self, lev, scal, s0, comp, flag, sg=None):
        a = [_mf(scal), _bs(s0)]
        self._check(self.fns["put_in_pert_form"](self.ctx, lev, *[C.byref(x) for x in a], comp,
                                                 int(flag),
                                                 C.byref(sg) if sg is not None else None),
                    "PutInPertForm")

    def MakeVelForce(self, lev, vel_force, uedge, rho, gpi, w0_force_cart, w0_cart, normal,
                     rho0, grav_cell, w0, buoyancy_cutoff_factor, base_cutoff_density,
                     do_add_utilde_force, average_base_state=False, sg=None):
        """Maestro::MakeVelForce on one level (without ROTATION)."""
        a = [_mf(vel_force), _face(uedge), _mf(rho), _mf(gpi), _mf(w0_force_cart), _mf(w0_cart),
             _mf(normal), _bs(rho0), _bs(grav_cell), _bs(w0)]
        self._check(self.fns["make_vel_force"](
            self.ctx, lev, *[C.byref(x) for x in a], float(buoyancy_cutoff_factor),
            float(base_cutoff_density), int(do_add_utilde_force), int(average_base_state),
            C.byref(sg) if sg is not None else None), "MakeVelForce")

    def ModifyScalForce(self, lev, scal_force, state, umac, s0_edge, s0_cart, w0_cart, w0, comp,
                        fullform, sg=None):
        """Maestro::ModifyScalForce on one level."""
        a = [_mf(scal_force), _mf(state), _face(umac), _bs(s0_edge), _mf(s0_cart), _mf(w0_cart),
             _bs(w0)]
        self._check(self.fns["modify_scal_force"](
            self.ctx, lev, *[C.byref(x) for x in a], comp, int(fullform),
            C.byref(sg) if sg is not None else None), "ModifyScalForce")

    def write_plotfile(self, path, mfs, varnames, time, level_steps, prob_lo, prob_hi):
        """amrex::WriteMultiLevelPlotfile as Maestro::WritePlotFile calls it
        (MaestroPlot.cpp:99-116): components 0..len(varnames)-1 of the valid
        cells of mfs[lev], lev = 0..len(mfs)-1."""
        nlev = len(mfs)
        descs = [_mf(m) for m in mfs]
        arr = (C.POINTER(_abi.MultiFabDesc) * nlev)(*[C.pointer(d) for d in descs])
        names = (C.c_char_p * len(varnames))(*[v.encode() for v in varnames])
        steps = (C.c_int * nlev)(*[int(s) for s in level_steps])
        lo = (C.c_double * 3)(*(list(prob_lo) + [0.0] * (3 - len(prob_lo))))
        hi = (C.c_double * 3)(*(list(prob_hi) + [0.0] * (3 - len(prob_hi))))
        self._check(self.fns["write_plotfile"](self.ctx, str(path).encode(), nlev, arr,
                                               len(varnames), names, float(time), steps, lo, hi),
                    "write_plotfile")

    def FillPatch(self, lev, mf, src, srccomp, destcomp, ncomp, startbccomp, bcs_in,
                  variable_type):
        """Maestro::FillPatch at level 0 (MaestroFillData.cpp:24-47): src is the
        MultiFab GetData picks for the fill time (mf_new or mf_old, possibly
        mf itself); bcs_in a BCRec array (bcs_s or bcs_u) or a list of
        (lo, hi) tuples."""
        if not isinstance(bcs_in, C.Array):
            bcs_in = bcrec_array(bcs_in)
        recs = bcrec_array(bcrec_to_list(bcs_in)[startbccomp:startbccomp + ncomp])
        if len(recs) != ncomp:
            raise _abi.MxError(_abi.MX_ERR_ARG, "FillPatch: bcs_in shorter than startbccomp + ncomp")
        a = [_mf(mf), _mf(src)]
        self._check(self.fns["fill_patch"](self.ctx, lev, C.byref(a[0]), C.byref(a[1]), srccomp,
                                           destcomp, ncomp, recs, int(variable_type)),
                    "FillPatch")

    def FillPatchTwoLevels(self, lev, mf, crse_src, fine_src, srccomp, destcomp, ncomp,
                           startbccomp, bcs_in, variable_type):
        """Maestro::FillPatch on a level lev > 0 (MaestroFillData.cpp:48-66):
        fine_src is level lev's and crse_src level lev - 1's MultiFab that
        GetData picks for the fill time (mf itself, possibly)."""
        if not isinstance(bcs_in, C.Array):
            bcs_in = bcrec_array(bcs_in)
        recs = bcrec_array(bcrec_to_list(bcs_in)[startbccomp:startbccomp + ncomp])
        if len(recs) != ncomp:
            raise _abi.MxError(_abi.MX_ERR_ARG, "FillPatch: bcs_in shorter than startbccomp + ncomp")
        a = [_mf(mf), _mf(crse_src), _mf(fine_src)]
        self._check(self.fns["fill_patch_two_levels"](self.ctx, lev, *[C.byref(x) for x in a],
                                                      srccomp, destcomp, ncomp, recs,
                                                      int(variable_type)), "FillPatchTwoLevels")

    def average_down(self, lev, fine, crse, comp, ncomp):
        """amrex::average_down between levels lev + 1 and lev (one step of
        Maestro::AverageDown)."""
        a = [_mf(fine), _mf(crse)]
        self._check(self.fns["average_down"](self.ctx, lev, *[C.byref(x) for x in a], comp, ncomp),
                    "average_down")

    def average_down_faces(self, lev, fine, crse):
        """amrex::average_down_faces between levels lev + 1 and lev."""
        a = [_face(fine), _face(crse)]
        self._check(self.fns["average_down_faces"](self.ctx, lev, *[C.byref(x) for x in a]),
                    "average_down_faces")

    def FillPatchUedge(self, uedge):
        """Maestro::FillPatchUedge (MaestroFillData.cpp:455-833): uedge[lev]
        is level lev's FaceMultiFab (one component, one ghost face)."""
        descs = [_face(f) for f in uedge]
        arr = (_abi.FaceMultiFabDesc * len(descs))(*descs)
        self._check(self.fns["fill_patch_uedge"](self.ctx, len(descs), arr), "fill_patch_uedge")

    # amrex::FluxRegister of level lev (Maestro::flux_reg_s[lev]), held by the
    # context
    def flux_reg_define(self, lev, ncomp):
        self._check(self.fns["flux_reg_define"](self.ctx, lev, ncomp), "FluxRegister")

    def flux_reg_setval(self, lev, v):
        self._check(self.fns["flux_reg_setval"](self.ctx, lev, float(v)), "FluxRegister::setVal")

    def flux_reg_crse_init(self, lev, flux, dir, srccomp, destcomp, numcomp, mult):
        """FluxRegister::CrseInit with level lev - 1's fluxes of direction
        dir (the MultiFab flux[dir] of a FaceMultiFab)."""
        m = _mf(flux)
        self._check(self.fns["flux_reg_crse_init"](self.ctx, lev, C.byref(m), dir, srccomp,
                                                   destcomp, numcomp, float(mult)),
                    "FluxRegister::CrseInit")

    def flux_reg_fine_add(self, lev, flux, dir, srccomp, destcomp, numcomp, mult):
        """FluxRegister::FineAdd with level lev's fluxes of direction dir."""
        m = _mf(flux)
        self._check(self.fns["flux_reg_fine_add"](self.ctx, lev, C.byref(m), dir, srccomp,
                                                  destcomp, numcomp, float(mult)),
                    "FluxRegister::FineAdd")

    def flux_reg_reflux(self, lev, state, scale, scomp, dcomp, ncomp):
        """FluxRegister::Reflux onto level lev - 1's state."""
        m = _mf(state)
        self._check(self.fns["flux_reg_reflux"](self.ctx, lev, C.byref(m), float(scale), scomp,
                                                dcomp, ncomp), "FluxRegister::Reflux")

    def flux_reg_read(self, lev):
        """Every register value of level lev as a numpy array (registers in
        (fine box, orientation) order, components outermost, i fastest)."""
        n = C.c_longlong(0)
        self._check(self.fns["flux_reg_read"](self.ctx, lev, None, 0, C.byref(n)), "flux_reg_read")
        out = np.zeros(n.value)
        self._check(self.fns["flux_reg_read"](self.ctx, lev,
                                              out.ctypes.data_as(C.POINTER(C.c_double)), n.value,
                                              C.byref(n)), "flux_reg_read")
        return out

    def AverageDown(self, mfs, comp, ncomp):
        """Maestro::AverageDown (MaestroFillData.cpp:145-154): mfs[lev] per level."""
        for lev in range(len(mfs) - 2, -1, -1):
            self.average_down(lev, mfs[lev + 1], mfs[lev], comp, ncomp)

    def AverageDownFaces(self, edges):
        """Maestro::AverageDownFaces (MaestroFillData.cpp:156-172)."""
        for lev in range(len(edges) - 2, -1, -1):
            self.average_down_faces(lev, edges[lev + 1], edges[lev])

    def ConvertRhoXToX(self, lev, scal, flag):
        m = _mf(scal)
        self._check(self.fns["convert_rhox_to_x"](self.ctx, lev, C.byref(m), int(flag)),
                    "ConvertRhoXToX")

    def ConvertRhoHToH(self, lev, scal, flag):
        m = _mf(scal)
        self._check(self.fns["convert_rhoh_to_h"](self.ctx, lev, C.byref(m), int(flag)),
                    "ConvertRhoHToH")

    def UpdateScal(self, lev, stateold, statenew, sflux, force, start_comp, num_comp,
                   p0_cart=None, base_cutoff_density=0.0, do_eos_h_above_cutoff=False):
        """Maestro::UpdateScal on one level (the valid cells; reflux, AverageDown
        and FillPatch stay with the caller)."""
        a = [_mf(stateold), _mf(statenew), _face(sflux), _mf(force)]
        p0 = _mf(p0_cart)
        self._check(self.fns["update_scal"](self.ctx, lev, *[C.byref(x) for x in a], start_comp,
                                            num_comp, C.byref(p0), float(base_cutoff_density),
                                            int(do_eos_h_above_cutoff)), "UpdateScal")

    def UpdateVel(self, lev, uold, unew, umac, uedge, force, sponge, w0mac, w0_cart,
                  do_sponge=False):
        """Maestro::UpdateVel on one level (the valid cells)."""
        a = [_mf(uold), _mf(unew), _face(umac), _face(uedge), _mf(force), _mf(sponge),
             _face(w0mac), _mf(w0_cart)]
        self._check(self.fns["update_vel"](self.ctx, lev, *[C.byref(x) for x in a],
                                           int(do_sponge)), "UpdateVel")

    def MakeRhoXFlux(self, lev, sflux, etarhoflux, sedge, umac, rho0_old, rho0_edge_old,
                     r0mac_old, rho0_new, rho0_edge_new, r0mac_new, rho0_predicted_edge, w0,
                     start_comp, num_comp):
        a = [_face(sflux), _mf(etarhoflux), _face(sedge), _face(umac), _bs(rho0_old),
             _bs(rho0_edge_old), _face(r0mac_old), _bs(rho0_new), _bs(rho0_edge_new),
             _face(r0mac_new), _bs(rho0_predicted_edge), _bs(w0)]
        self._check(self.fns["make_rhox_flux"](self.ctx, lev, *[C.byref(x) for x in a],
                                               start_comp, num_comp), "MakeRhoXFlux")

    def MakeRhoHFlux(self, lev, sflux, sedge, umac, w0mac, rho0_old, rho0_edge_old, r0mac_old,
                     rho0_new, rho0_edge_new, r0mac_new, rhoh0_old, rhoh0_edge_old, rh0mac_old,
                     rhoh0_new, rhoh0_edge_new, rh0mac_new, h0mac_old, h0mac_new):
        a = [_face(sflux), _face(sedge), _face(umac), _face(w0mac), _bs(rho0_old),
             _bs(rho0_edge_old), _face(r0mac_old), _bs(rho0_new), _bs(rho0_edge_new),
             _face(r0mac_new), _bs(rhoh0_old), _bs(rhoh0_edge_old), _face(rh0mac_old),
             _bs(rhoh0_new), _bs(rhoh0_edge_new), _face(rh0mac_new), _face(h0mac_old),
             _face(h0mac_new)]
        self._check(self.fns["make_rhoh_flux"](self.ctx, lev, *[C.byref(x) for x in a]),
                    "MakeRhoHFlux")

    def CelltoEdge(self, lev, nr, s0_cell, s0_edge):
        a = [_bs(s0_cell), _bs(s0_edge)]
        self._check(self.fns["cell_to_edge"](self.ctx, lev, nr, *[C.byref(x) for x in a]),
                    "CelltoEdge")

    def MakeEdgeState1d(self, rg, s, sedge, force, w0):
        """Maestro::MakeEdgeState1d; rg is an _abi.RadialGeom, w0 the Maestro
        member w0 (edge-centred)."""
        a = [_bs(s), _bs(sedge), _bs(force), _bs(w0)]
        self._check(self.fns["make_edge_state_1d"](self.ctx, C.byref(rg),
                                                   *[C.byref(x) for x in a]),
                    "MakeEdgeState1d")

    @staticmethod
    def pass_args(f, keep):
        """PassArgs of the fields of dict `f` (descriptors kept alive in `keep`)."""
        a = _abi.PassArgs()
        for name in ("utilde", "ufull", "vel_force", "w0_cart", "scal", "scal_force", "etarhoflux"):
            d = _mf(f[name])
            keep.append(d)
            setattr(a, name, d)
        for name in ("w0mac", "r0mac", "rh0mac", "h0mac", "utrans", "umac", "sedge", "sflux"):
            d = _face(f.get(name))
            keep.append(d)
            setattr(a, name, d)
        for name in ("w0", "rho0", "rho0_edge", "rhoh0", "rhoh0_edge", "rho0_predicted_edge"):
            setattr(a, name, _bs(f[name]))
        return a

    def advance_pass(self, lev, f):
        """mx_adv_advance_pass with the fields of dict `f` (see Workload.pass_fields)."""
        keep = []
        a = self.pass_args(f, keep)
        self._check(self.fns["advance_pass"](self.ctx, lev, C.byref(a)), "advance_pass")

    def advance_pass_noalloc(self, lev, f):
        """mx_adv_advance_pass_noalloc: the same pass with its tables and
        workspace in the library's fixed device scratch (no allocation)."""
        keep = []
        a = self.pass_args(f, keep)
        self._check(self.fns["advance_pass_noalloc"](self.ctx, lev, C.byref(a)),
                    "advance_pass_noalloc")

    def _levels_args(self, fields):
        keep = []
        arr = (_abi.PassArgs * len(fields))(*[self.pass_args(f, keep) for f in fields])
        return arr, keep

    def advance_pass_levels(self, fields):
        """The pass over levels 0..len(fields)-1 (fields[lev] per level)."""
        arr, keep = self._levels_args(fields)
        self._check(self.fns["advance_pass_levels"](self.ctx, len(fields), arr),
                    "advance_pass_levels")

    def pass_graph(self, fields):
        """Capture the hierarchy pass as a HIP graph; returns a PassGraph."""
        arr, keep = self._levels_args(fields)
        g = C.c_void_p()
        self._check(self.fns["pass_graph_create"](self.ctx, len(fields), arr, C.byref(g)),
                    "pass_graph_create")
        return PassGraph(self, g, keep)

    def autotune(self, fields):
        """Pick the fused edge kernel's tile shape and z-chunk; returns the
        candidates as a list of (tile, tz, ms) and the chosen index."""
        arr, keep = self._levels_args(fields)
        rep = _abi.TuneReport()
        self._check(self.fns["autotune"](self.ctx, len(fields), arr, C.byref(rep)), "autotune")
        cands = [(rep.tile[i], rep.tz[i], rep.ms[i]) for i in range(rep.ncand)]
        return cands, rep.best


class PassGraph:
    """A captured hierarchy pass (mx_adv_pass_graph_*)."""

    def __init__(self, m, handle, keep):
        self.m = m
        self.handle = handle
        self._keep = keep

    def launch(self):
        self.m._check(self.m.fns["pass_graph_launch"](self.handle), "pass_graph_launch")

    def close(self):
        if self.handle:
            self.m.fns["pass_graph_destroy"](self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bcrec_array(recs):
    """list of ((lo0,lo1,lo2),(hi0,hi1,hi2)) -> ctypes BCRec array"""
    arr = (_abi.BCRec * len(recs))()
    for n, (lo, hi) in enumerate(recs):
        for d in range(3):
            arr[n].lo[d] = lo[d]
            arr[n].hi[d] = hi[d]
    return arr


def bcrec_to_list(arr):
    return [(tuple(r.lo), tuple(r.hi)) for r in arr]


def max_abs(a):
    return float(np.max(np.abs(a))) if a.size else 0.0
