"""ctypes mirror of the C-ABI types in include/mx_adv.h.

The structs here are byte-for-byte the ones the HIP library
(libmaestroex_amd.so) takes; `bind()` attaches argument/return types to the
exported `<prefix>*` functions of a loaded library.
"""
import ctypes as C

MX_MAX_LEV = 8
MX_MAX_SCAL = 32

MX_OK = 0
MX_ERR_ARG = -1
MX_ERR_BOX = -2
MX_ERR_BC = -3
MX_ERR_HIP = -4
MX_ERR_UNSUPPORTED = -5
MX_ERR_NOMEM = -6
MX_ERR_NOLIB = -7
MX_ERR_IO = -8
MX_ERR_SCRATCH = -9

# amrex::BCType
BC_REFLECT_ODD, BC_INT_DIR, BC_REFLECT_EVEN, BC_FOEXTRAP, BC_EXT_DIR, BC_HOEXTRAP = -1, 0, 1, 2, 3, 4
# amrex::PhysBCType
PHYS_INTERIOR, PHYS_INFLOW, PHYS_OUTFLOW, PHYS_SYMMETRY, PHYS_SLIPWALL, PHYS_NOSLIPWALL = 0, 1, 2, 3, 4, 5

ERRORS = {
    MX_ERR_ARG: "bad argument",
    MX_ERR_BOX: "FAB does not cover the region",
    MX_ERR_BC: "invalid boundary type",
    MX_ERR_HIP: "HIP runtime failure",
    MX_ERR_UNSUPPORTED: "unsupported configuration",
    MX_ERR_NOMEM: "workspace allocation failed",
    MX_ERR_NOLIB: "no usable GPU",
    MX_ERR_IO: "file could not be created or written",
    MX_ERR_SCRATCH: "needs more device memory than the no-alloc scratch holds",
}


class MxError(RuntimeError):
    def __init__(self, code, where=""):
        self.code = code
        super().__init__(f"{where}: error {code} ({ERRORS.get(code, 'unknown')})")


class Box(C.Structure):
    _fields_ = [("lo", C.c_int * 3), ("hi", C.c_int * 3)]


class Fab(C.Structure):
    _fields_ = [
        ("p", C.c_void_p),
        ("jstride", C.c_longlong),
        ("kstride", C.c_longlong),
        ("nstride", C.c_longlong),
        ("begin", C.c_int * 3),
        ("end", C.c_int * 3),
        ("ncomp", C.c_int),
    ]


class MultiFabDesc(C.Structure):
    _fields_ = [("nfab", C.c_int), ("fab", C.POINTER(Fab))]


class FaceMultiFabDesc(C.Structure):
    _fields_ = [("dir", MultiFabDesc * 3)]


class BaseStateDesc(C.Structure):
    _fields_ = [("p", C.c_void_p), ("nlev", C.c_int), ("len", C.c_int)]


class BCRec(C.Structure):
    _fields_ = [("lo", C.c_int * 3), ("hi", C.c_int * 3)]


class Params(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("ppm_type", C.c_int),
        ("slope_order", C.c_int),
        ("ppm_trace_forces", C.c_int),
        ("species_pred_type", C.c_int),
        ("enthalpy_pred_type", C.c_int),
        ("spherical", C.c_int),
        ("use_exact_base_state", C.c_int),
        ("evolve_base_state", C.c_int),
        ("dt", C.c_double),
        ("rel_eps", C.c_double),
        ("finest_level", C.c_int),
        ("domain_lo", (C.c_int * 3) * MX_MAX_LEV),
        ("domain_hi", (C.c_int * 3) * MX_MAX_LEV),
        ("dx", (C.c_double * 3) * MX_MAX_LEV),
        ("is_periodic", C.c_int * 3),
        ("phys_bc_lo", C.c_int * 3),
        ("phys_bc_hi", C.c_int * 3),
        ("nscal", C.c_int),
        ("comp_rho", C.c_int),
        ("comp_rhoh", C.c_int),
        ("comp_temp", C.c_int),
        ("comp_pi", C.c_int),
        ("first_spec", C.c_int),
        ("num_spec", C.c_int),
        ("reflux_type", C.c_int),
    ]


MX_MAX_CHUNK = 16


class SphGeom(C.Structure):
    """mx_sph_geom: spherical geometry of the base-state producers."""
    _fields_ = [
        ("center", C.c_double * 3),
        ("prob_lo", C.c_double * 3),
        ("dr_fine", C.c_double),
        ("nr_fine", C.c_int),
        ("r_cc_loc", C.c_void_p),
        ("r_edge_loc", C.c_void_p),
        ("s0_interp_type", C.c_int),
        ("w0_interp_type", C.c_int),
        ("s0mac_interp_type", C.c_int),
        ("w0mac_interp_type", C.c_int),
    ]


class RadialGeom(C.Structure):
    """mx_radial_geom: radial levels, cell counts, widths and chunks."""
    _fields_ = [
        ("nlev", C.c_int),
        ("nr_fine", C.c_int),
        ("nr", C.c_int * MX_MAX_LEV),
        ("dr", C.c_double * MX_MAX_LEV),
        ("nchunks", C.c_int * MX_MAX_LEV),
        ("r_start", (C.c_int * MX_MAX_CHUNK) * MX_MAX_LEV),
        ("r_end", (C.c_int * MX_MAX_CHUNK) * MX_MAX_LEV),
    ]


def radial_geom(nr_fine, dr, levels=None):
    """One radial level covering [0, nr_fine) (levels=None), or the given
    levels: [(nr, dr, [(r_start, r_end), ...]), ...]."""
    g = RadialGeom()
    g.nr_fine = nr_fine
    if levels is None:
        levels = [(nr_fine, dr, [(0, nr_fine - 1)])]
    g.nlev = len(levels)
    for n, (nr, d, chunks) in enumerate(levels):
        g.nr[n] = nr
        g.dr[n] = d
        g.nchunks[n] = len(chunks)
        for i, (a, b) in enumerate(chunks):
            g.r_start[n][i] = a
            g.r_end[n][i] = b
    return g


class PassArgs(C.Structure):
    _fields_ = [
        ("utilde", MultiFabDesc),
        ("ufull", MultiFabDesc),
        ("vel_force", MultiFabDesc),
        ("w0_cart", MultiFabDesc),
        ("scal", MultiFabDesc),
        ("scal_force", MultiFabDesc),
        ("w0mac", FaceMultiFabDesc),
        ("r0mac", FaceMultiFabDesc),
        ("rh0mac", FaceMultiFabDesc),
        ("h0mac", FaceMultiFabDesc),
        ("w0", BaseStateDesc),
        ("rho0", BaseStateDesc),
        ("rho0_edge", BaseStateDesc),
        ("rhoh0", BaseStateDesc),
        ("rhoh0_edge", BaseStateDesc),
        ("rho0_predicted_edge", BaseStateDesc),
        ("utrans", FaceMultiFabDesc),
        ("umac", FaceMultiFabDesc),
        ("sedge", FaceMultiFabDesc),
        ("sflux", FaceMultiFabDesc),
        ("etarhoflux", MultiFabDesc),
    ]


P = C.POINTER
_vp = C.c_void_p
_i = C.c_int
_d = C.c_double

# name -> argtypes (after the prefix)
SIGNATURES = {
    "create": [P(Params), P(_vp)],
    "destroy": [_vp],
    "set_level": [_vp, _i, _i, P(Box)],
    "set_dt": [_vp, _d, _d],
    "bc_setup": [_vp, P(BCRec), P(BCRec)],
    "ppm": [_vp, _i, P(Box), P(Fab), P(Fab), P(Fab), P(Fab), P(Fab), P(Fab), P(BCRec), _i, _i, _i, _i],
    "slope": [_vp, _i, _i, P(Box), P(Fab), P(Fab), P(BCRec), _i, _i, _i, _i],
    "make_utrans": [_vp, _i, P(MultiFabDesc), P(MultiFabDesc), P(FaceMultiFabDesc), P(FaceMultiFabDesc), P(MultiFabDesc)],
    "addw0": [_vp, _i, P(FaceMultiFabDesc), P(FaceMultiFabDesc), P(BaseStateDesc), _d],
    "fill_face_ghosts": [_vp, _i, P(FaceMultiFabDesc)],
    "velpred": [_vp, _i, P(MultiFabDesc), P(MultiFabDesc), P(FaceMultiFabDesc), P(FaceMultiFabDesc),
                P(FaceMultiFabDesc), P(MultiFabDesc), P(MultiFabDesc)],
    "make_edge_scal": [_vp, _i, P(MultiFabDesc), P(FaceMultiFabDesc), P(FaceMultiFabDesc), P(MultiFabDesc),
                       _i, P(BCRec), _i, _i, _i, _i, _i],
    "make_rhox_flux": [_vp, _i, P(FaceMultiFabDesc), P(MultiFabDesc), P(FaceMultiFabDesc), P(FaceMultiFabDesc),
                       P(BaseStateDesc), P(BaseStateDesc), P(FaceMultiFabDesc), P(BaseStateDesc),
                       P(BaseStateDesc), P(FaceMultiFabDesc), P(BaseStateDesc), P(BaseStateDesc), _i, _i],
    "make_rhoh_flux": [_vp, _i, P(FaceMultiFabDesc), P(FaceMultiFabDesc), P(FaceMultiFabDesc),
                       P(FaceMultiFabDesc), P(BaseStateDesc), P(BaseStateDesc), P(FaceMultiFabDesc),
                       P(BaseStateDesc), P(BaseStateDesc), P(FaceMultiFabDesc), P(BaseStateDesc),
                       P(BaseStateDesc), P(FaceMultiFabDesc), P(BaseStateDesc), P(BaseStateDesc),
                       P(FaceMultiFabDesc), P(FaceMultiFabDesc), P(FaceMultiFabDesc)],
    "cell_to_edge": [_vp, _i, _i, P(BaseStateDesc), P(BaseStateDesc)],
    "make_edge_state_1d": [_vp, P(RadialGeom), P(BaseStateDesc), P(BaseStateDesc),
                           P(BaseStateDesc), P(BaseStateDesc)],
    "advance_pass": [_vp, _i, P(PassArgs)],
    "put_1d_array_on_cart": [_vp, _i, P(BaseStateDesc), P(MultiFabDesc), _i, _i, P(SphGeom)],
    "make_w0mac": [_vp, _i, P(BaseStateDesc), P(MultiFabDesc), P(FaceMultiFabDesc), P(SphGeom)],
    "make_s0mac": [_vp, _i, P(BaseStateDesc), P(MultiFabDesc), P(FaceMultiFabDesc), P(SphGeom)],
    "put_in_pert_form": [_vp, _i, P(MultiFabDesc), P(BaseStateDesc), _i, _i, P(SphGeom)],
    "make_vel_force": [_vp, _i, P(MultiFabDesc), P(FaceMultiFabDesc), P(MultiFabDesc), P(MultiFabDesc),
                       P(MultiFabDesc), P(MultiFabDesc), P(MultiFabDesc), P(BaseStateDesc),
                       P(BaseStateDesc), P(BaseStateDesc), _d, _d, _i, _i, P(SphGeom)],
    "modify_scal_force": [_vp, _i, P(MultiFabDesc), P(MultiFabDesc), P(FaceMultiFabDesc),
                          P(BaseStateDesc), P(MultiFabDesc), P(MultiFabDesc), P(BaseStateDesc), _i,
                          _i, P(SphGeom)],
    "write_plotfile": [_vp, C.c_char_p, _i, P(P(MultiFabDesc)), _i, P(C.c_char_p), _d, P(_i),
                       P(_d), P(_d)],
    "fill_patch": [_vp, _i, P(MultiFabDesc), P(MultiFabDesc), _i, _i, _i, P(BCRec), _i],
    "fill_patch_two_levels": [_vp, _i, P(MultiFabDesc), P(MultiFabDesc), P(MultiFabDesc), _i, _i,
                              _i, P(BCRec), _i],
    "average_down": [_vp, _i, P(MultiFabDesc), P(MultiFabDesc), _i, _i],
    "average_down_faces": [_vp, _i, P(FaceMultiFabDesc), P(FaceMultiFabDesc)],
    "fill_patch_uedge": [_vp, _i, P(FaceMultiFabDesc)],
    "flux_reg_define": [_vp, _i, _i],
    "flux_reg_setval": [_vp, _i, _d],
    "flux_reg_crse_init": [_vp, _i, P(MultiFabDesc), _i, _i, _i, _i, _d],
    "flux_reg_fine_add": [_vp, _i, P(MultiFabDesc), _i, _i, _i, _i, _d],
    "flux_reg_reflux": [_vp, _i, P(MultiFabDesc), _d, _i, _i, _i],
    "flux_reg_read": [_vp, _i, P(_d), C.c_longlong, P(C.c_longlong)],
    "convert_rhox_to_x": [_vp, _i, P(MultiFabDesc), _i],
    "convert_rhoh_to_h": [_vp, _i, P(MultiFabDesc), _i],
    "update_scal": [_vp, _i, P(MultiFabDesc), P(MultiFabDesc), P(FaceMultiFabDesc), P(MultiFabDesc),
                    _i, _i, P(MultiFabDesc), _d, _i],
    "update_vel": [_vp, _i, P(MultiFabDesc), P(MultiFabDesc), P(FaceMultiFabDesc),
                   P(FaceMultiFabDesc), P(MultiFabDesc), P(MultiFabDesc), P(FaceMultiFabDesc),
                   P(MultiFabDesc), _i],
}

MX_TUNE_MAX = 16


class TuneReport(C.Structure):
    _fields_ = [("ncand", C.c_int), ("tile", C.c_int * MX_TUNE_MAX), ("tz", C.c_int * MX_TUNE_MAX),
                ("ms", C.c_double * MX_TUNE_MAX), ("best", C.c_int)]


# entry points only the HIP library has
LIB_ONLY = {
    "advance_pass_levels": [_vp, _i, P(PassArgs)],
    "advance_pass_noalloc": [_vp, _i, P(PassArgs)],
    "pass_graph_create": [_vp, _i, P(PassArgs), P(_vp)],
    "pass_graph_launch": [_vp],
    "pass_graph_destroy": [_vp],
    "autotune": [_vp, _i, P(PassArgs), P(TuneReport)],
    "scal_step": [_vp, P(Fab), P(Fab), P(Fab), _i, _i],
    "scal_step_vjp": [_vp, P(Fab), P(Fab), P(Fab), P(Fab), P(Fab), _i, _i],
    "scal_step_boxes": [_vp, P(MultiFabDesc), P(FaceMultiFabDesc), P(MultiFabDesc), _i, _i],
    "scal_step_vjp_boxes": [_vp, P(MultiFabDesc), P(FaceMultiFabDesc), P(MultiFabDesc),
                            P(MultiFabDesc), P(FaceMultiFabDesc), _i, _i],
    "set_stream": [_vp, _vp],
    "sync": [_vp],
    "error_string": [_i],
    "prof_enable": [_vp, _i],
    "prof_read": [_vp, _i, P(C.c_char_p), P(C.c_double), P(C.c_longlong)],
    "set_tuning": [_vp, C.c_char_p, C.c_longlong],
}


# entry points with no counterpart in the C oracle library (the plotfile
# writer is checked against a file-level restatement instead)
NO_ORACLE = {"write_plotfile"}


def bind(lib, prefix, names=None, extra=None):
    """Attach ctypes signatures to `lib.<prefix><name>`; returns {name: fn}."""
    table = dict(SIGNATURES)
    if extra:
        table.update(extra)
    out = {}
    for name, argt in table.items():
        if names is not None and name not in names:
            continue
        fn = getattr(lib, prefix + name)
        fn.argtypes = argt
        fn.restype = C.c_char_p if name == "error_string" else C.c_int
        out[name] = fn
    return out
