/*
 * mx_adv.h -- C-ABI of the MI355X advection hot path (MAESTROeX-compatible).
 *
 * The path: PPM / slope reconstruction, transverse-corrected edge-state
 * prediction (MakeUtrans, VelPred, MakeEdgeScal) and the conservative flux
 * construction (MakeRhoXFlux, MakeRhoHFlux) with the base-state w0 / rho0
 * terms.  Every entry point below replaces one member function of
 * `class Maestro` (Source/Maestro.H in the reference); each declaration cites
 * the function it replaces.
 *
 * Conventions
 *   - All field data are FP64 and live where the caller put them: device
 *     memory for the HIP library (libmaestroex_amd.so).  The caller owns and
 *     pre-allocates every input and output, exactly as the reference callers
 *     do (MaestroAdvection.cpp:41-55, MaestroAdvance.cpp:167-192).
 *   - A FAB is described by an Array4-like view (mx_fab): element (i,j,k,n) is
 *     p[(i-begin[0]) + (j-begin[1])*jstride + (k-begin[2])*kstride + n*nstride]
 *     with i fastest, component outermost (AMReX Array4 order).  2-D data use
 *     k == 0 and begin[2] == 0, end[2] == 1.
 *   - A level's MultiFab is one mx_fab per box, in the order of the boxes
 *     given to mx_adv_set_level (the level's BoxArray).  Face-centred data
 *     (MAC velocities, edge states, fluxes) are one MultiFab per direction.
 *   - One-dimensional base states (BaseState<Real>, Source/BaseState.H:80) are
 *     (lev, r) arrays: element (lev, r) is p[lev*len + r].
 *   - Ghost cells of inputs are filled by the caller, as the reference
 *     assumes.  Functions that fill ghost faces in the reference
 *     (MakeUtrans, Addw0) do so here too (single-level FillBoundary with
 *     periodicity, then FillUmacGhost).
 *   - Calls are asynchronous on the context's HIP stream; mx_adv_sync waits.
 *   - Every function returns MX_OK (0) or a negative error code; no call
 *     aborts the process (the reference calls amrex::Abort / Error instead).
 */
#ifndef MX_ADV_H
#define MX_ADV_H

#ifdef __cplusplus
extern "C" {
#endif

#define MX_MAX_LEV 8
#define MX_MAX_SCAL 32

/* return codes */
enum {
    MX_OK = 0,
    MX_ERR_ARG = -1,         /* null pointer, bad count, bad parameter value */
    MX_ERR_BOX = -2,         /* a FAB does not cover the region the stage needs */
    MX_ERR_BC = -3,          /* invalid boundary type (VelPred's "invalid boundary type") */
    MX_ERR_HIP = -4,         /* HIP runtime failure */
    MX_ERR_UNSUPPORTED = -5, /* configuration the path does not implement */
    MX_ERR_NOMEM = -6,       /* workspace allocation failed */
    MX_ERR_NOLIB = -7,       /* no usable GPU */
    MX_ERR_IO = -8,          /* a file could not be created or written */
    MX_ERR_SCRATCH = -9      /* the call needs more device memory than the fixed
                                scratch of mx_adv_advance_pass_noalloc holds */
};

/* amrex::BCType values (mathematical boundary types, per component) */
enum {
    MX_BC_REFLECT_ODD = -1,
    MX_BC_INT_DIR = 0,
    MX_BC_REFLECT_EVEN = 1,
    MX_BC_FOEXTRAP = 2,
    MX_BC_EXT_DIR = 3,
    MX_BC_HOEXTRAP = 4
};

/* amrex::PhysBCType values (physical boundary types, per face) */
enum {
    MX_PHYS_INTERIOR = 0,
    MX_PHYS_INFLOW = 1,
    MX_PHYS_OUTFLOW = 2,
    MX_PHYS_SYMMETRY = 3,
    MX_PHYS_SLIPWALL = 4,
    MX_PHYS_NOSLIPWALL = 5
};

/* cell-centred index box, inclusive bounds (amrex::Box of a BoxArray) */
typedef struct {
    int lo[3];
    int hi[3];
} mx_box;

/* Array4-like view of one FAB */
typedef struct {
    double* p;
    long long jstride;
    long long kstride;
    long long nstride;
    int begin[3];
    int end[3]; /* exclusive */
    int ncomp;
} mx_fab;

/* one level of a MultiFab: fab[b] belongs to box b of the level */
typedef struct {
    int nfab;
    const mx_fab* fab;
} mx_multifab;

/* face-centred MultiFab triple (std::array<MultiFab, AMREX_SPACEDIM>) */
typedef struct {
    mx_multifab dir[3];
} mx_face_multifab;

/* BaseState<Real> with one variable: element (lev, r) is p[lev*len + r] */
typedef struct {
    double* p;
    int nlev;
    int len;
} mx_basestate;

/* amrex::BCRec */
typedef struct {
    int lo[3];
    int hi[3];
} mx_bcrec;

/* The Maestro members the path reads implicitly (Maestro.H; defaults from
   Source/param/_cpp_parameters). */
typedef struct {
    int dim;                 /* AMREX_SPACEDIM: 2 or 3 */
    int ppm_type;            /* 0 (slopes), 1 (CW84), 2 (CS08) */
    int slope_order;         /* 0, 2 or 4; used when ppm_type == 0 */
    int ppm_trace_forces;    /* 0 or 1 */
    int species_pred_type;   /* 1 rhoprime_and_X, 2 rhoX, 3 rho_and_X */
    int enthalpy_pred_type;  /* 0 rhoh, 1 rhohprime, 2 h, 3 T_then_rhohprime,
                                4 T_then_h, 5 hprime, 6 Tprime_then_h */
    int spherical;           /* 0 planar, 1 spherical (3-D only) */
    int use_exact_base_state;
    int evolve_base_state;
    double dt;
    double rel_eps;          /* 1e-8 * umax (MaestroDt.cpp:481-483) */
    int finest_level;        /* levels 0..finest_level */
    int domain_lo[MX_MAX_LEV][3];
    int domain_hi[MX_MAX_LEV][3];
    double dx[MX_MAX_LEV][3];
    int is_periodic[3];
    int phys_bc_lo[3];       /* phys_bc (Maestro.H:1917) */
    int phys_bc_hi[3];
    int nscal;               /* Nscal (Source/param/_variables:8-14) */
    int comp_rho;            /* Rho */
    int comp_rhoh;           /* RhoH */
    int comp_temp;           /* Temp */
    int comp_pi;             /* Pi */
    int first_spec;          /* FirstSpec */
    int num_spec;            /* NumSpec; FirstAux.. are the remaining comps */
    int reflux_type;         /* 0 none, 1 (default) AverageDownFaces of the fluxes,
                                2 flux registers (_cpp_parameters:24) */
} mx_adv_params;

typedef struct mx_adv_ctx mx_adv_ctx;

/* ---- lifecycle ---------------------------------------------------------- */

/* Create a context holding the Maestro members above.  Takes no reference
   counterpart beyond Maestro::Setup / ReadParameters (MaestroSetup.cpp). */
int mx_adv_create(const mx_adv_params* params, mx_adv_ctx** out);
int mx_adv_destroy(mx_adv_ctx* ctx);

/* grids[lev]: the level's BoxArray (Maestro.H, AmrCore::grids): disjoint
   boxes inside the level's domain (MX_ERR_BOX otherwise). */
int mx_adv_set_level(mx_adv_ctx* ctx, int lev, int nbox, const mx_box* boxes);

/* dt and rel_eps, set per step in the reference by EstDt (MaestroDt.cpp:481-483). */
int mx_adv_set_dt(mx_adv_ctx* ctx, double dt, double rel_eps);

/* Maestro::BCSetup (MaestroSetup.cpp:276-559): the per-component BCRec
   tables bcs_s (nscal entries) and bcs_u (dim entries) implied by phys_bc.
   The context uses the same tables inside MakeUtrans, VelPred and the pass. */
int mx_adv_bc_setup(const mx_adv_ctx* ctx, mx_bcrec* bcs_s, mx_bcrec* bcs_u);

/* Stream for subsequent calls (hipStream_t, NULL = default stream). */
int mx_adv_set_stream(mx_adv_ctx* ctx, void* hip_stream);
int mx_adv_sync(mx_adv_ctx* ctx);
const char* mx_adv_error_string(int code);

/* Kernel timing (instrumentation; no counterpart in the reference, whose
   timers are BL_PROFILE_VAR regions).  mx_adv_prof_enable(ctx, 1) clears the
   counters and brackets every later kernel launch with a HIP event pair on
   the context stream; mx_adv_prof_read(ctx, -1, ...) synchronises, folds the
   events into per-kernel totals and returns the number of kernel names;
   mx_adv_prof_read(ctx, i, &name, &ms, &n) returns kernel i's total time in
   ms and launch count. */
int mx_adv_prof_enable(mx_adv_ctx* ctx, int on);
int mx_adv_prof_read(mx_adv_ctx* ctx, int idx, const char** name, double* total_ms,
                     long long* launches);

/* Kernel selection and launch shape (no counterpart in the reference, which
   has one CPU loop nest per stage).  Every choice computes bit-identical
   results; they differ in speed only.  Keys:
     "edge_scal_fused"  1 (default): 3-D MakeEdgeScal with ppm_type 1/2 and
                        no traced forces runs as one z-marching kernel;
                        0: the three-kernel form (PPM, transverse, edges),
                        which also serves ppm_type 0, traced forces and 2-D.
     "edge_scal_tz"     planes per z-chunk of the fused kernel (default 32).
     "edge_scal_tile"   cell columns per block of the fused kernel, and the
                        block size: 0: 32 x 8 (384 threads), 1: 16 x 16 (384),
                        2: 16 x 12 (256), 3: 16 x 8 (192), 4: 16 x 11 (256);
                        -1 (default): 16 x 11.
     "utrans_line"      1 (default): 3-D MakeUtrans with ppm_type 1/2 traces
                        each cell once (x-faces from rows of cells through
                        LDS, y/z-faces by columns marching along y or z);
                        0: one face per thread, each cell traced for both
                        of its faces.
     "utrans_chunks"    chunks per y/z column of that form (1-64, default 8).
     "velpred_ppm_all"  1 (default): VelPred's PPM kernel computes every
                        velocity component of a cell in one thread (3-D,
                        ppm_type 1/2, no traced forces); 0: one component
                        per thread.
     "velpred_trans_all" 1 (default): VelPred's transverse kernel computes
                        the six transverse states of a cell in one thread
                        (3-D); 0: one (direction, correction) pair per thread.
     "velpred_fused"    1 (default): 3-D VelPred with ppm_type 1/2 and no
                        traced forces runs as one z-marching kernel (PPM,
                        interface states, transverse and final states on
                        chip); 0: the staged kernels.
     "velpred_fused_tz" planes per z-chunk of the fused VelPred (default 32).
     "velpred_fused_comp" 1 (default): one velocity component per block of
                        the fused VelPred; 0: all three in one block.
     "velpred_fused_tile" tile and register budget of the per-component
                        fused VelPred: 0: 16 x 11 columns, 3 waves per SIMD;
                        1: 16 x 8, 3; 2 (default): 16 x 11, 4; 3: 32 x 8, 3;
                        4: 16 x 8, 4; 5: 16 x 12, 4 (2, 4 and 5: tiles at
                        an acting x/y domain side at 3).
     "rhoh_flux_epilogue" 1 (default): in the pass, MakeRhoHFlux's planar
                        branches that read only the RhoH edge state
                        (enthalpy_pred_type 0, 1, 3) run as the fused edge
                        kernel's epilogue; 0: the flux kernel.
     "xcd_remap"        1: the staged VelPred, MakeUtrans and flux kernels
                        give each of the 8 XCDs one contiguous eighth of a
                        box's blocks; 0 (default): blocks in launch order.
     "edge_scal_xcd", "velpred_fused_xcd"  block order of the fused edge
                        kernel and of the per-component fused VelPred: 1
                        (the edge kernel's default): each XCD runs the
                        component blocks of a tile back to back (they share
                        the velocity planes they stage in that XCD's L2),
                        tiles round robin over the XCDs; 2: the same with a
                        contiguous eighth of the tiles per XCD; 0: tile x
                        component in launch order; -1 (VelPred's default,
                        VelPred only): 2 for one level of equal boxes, else
                        1.
     "adjoint_face_sums", "adjoint_state_fused", "adjoint_pow2_index",
     "adjoint_trace_rolled", "adjoint_xcd"  arrangements of the step
                        adjoint's kernels (defaults 0, 1, 1, 0, 0); every
                        choice gives bit-identical results.
   Unknown keys and out-of-range values return MX_ERR_ARG. */
int mx_adv_set_tuning(mx_adv_ctx* ctx, const char* key, long long value);

/* ---- unit-level kernels ------------------------------------------------- */

/* Maestro::PPM (Maestro.H:1396-1406, MaestroPPM.cpp:6-2263).  bx is the box
   Ip/Im are computed on; u,v,w are face velocities (is_umac != 0) or
   cell-centred one-component tracing velocities. */
int mx_adv_ppm(mx_adv_ctx* ctx, int lev, const mx_box* bx, const mx_fab* s,
               const mx_fab* u, const mx_fab* v, const mx_fab* w, mx_fab* Ip,
               mx_fab* Im, const mx_bcrec* bcs, int nbcs, int is_umac,
               int comp, int bccomp);

/* Maestro::Slopex / Slopey / Slopez (Maestro.H:1576-1593,
   MaestroSlopes.cpp:6-600); dir selects the function. */
int mx_adv_slope(mx_adv_ctx* ctx, int lev, int dir, const mx_box* bx,
                 const mx_fab* s, mx_fab* sl, const mx_bcrec* bcs, int nbcs,
                 int ncomp, int bc_start_comp, int slope_order);

/* ---- level-wide stages -------------------------------------------------- */
/* Every stage works on one level.  Where the reference fills ghost faces
   (MakeUtrans, Addw0), this library does what the reference does when
   finest_level == 0 (FillBoundary within the level, periodic images
   included, then FillUmacGhost at the physical boundaries of the level's
   domain) on every level.  With finest_level > 0 the reference instead
   fills coarse-fine ghost faces from the coarser level (FillPatch,
   MaestroMakeUtrans.cpp:641-647); here those faces keep the values the
   caller's arrays hold. */

/* Maestro::MakeUtrans (Maestro.H:1172-1177, MaestroMakeUtrans.cpp:6-648),
   including its closing FillBoundary + FillUmacGhost.  w0_cart is the
   cell-centred w0 (Maestro.H:1903) with at least one ghost cell: the
   vertical Riemann solve reads it at the face's index, the top face's
   included, as the reference does; w0mac is read only when spherical. */
int mx_adv_make_utrans(mx_adv_ctx* ctx, int lev, const mx_multifab* utilde,
                       const mx_multifab* ufull, mx_face_multifab* utrans,
                       const mx_face_multifab* w0mac,
                       const mx_multifab* w0_cart);

/* Maestro::Addw0 (Maestro.H:693-696, MaestroFill3dData.cpp:357-461),
   including its closing FillBoundary + FillUmacGhost. */
int mx_adv_addw0(mx_adv_ctx* ctx, int lev, mx_face_multifab* uedge,
                 const mx_face_multifab* w0mac, const mx_basestate* w0,
                 double mult);

/* Maestro::VelPred (Maestro.H:1711-1718, MaestroVelPred.cpp:5-292).
   w0_cart as for MakeUtrans (one ghost cell). */
int mx_adv_velpred(mx_adv_ctx* ctx, int lev, const mx_multifab* utilde,
                   const mx_multifab* ufull, const mx_face_multifab* utrans,
                   mx_face_multifab* umac, const mx_face_multifab* w0mac,
                   const mx_multifab* force, const mx_multifab* w0_cart);

/* Maestro::MakeEdgeScal (Maestro.H:913-919,
   MaestroMakeEdgeScalars.cpp:6-302). */
int mx_adv_make_edge_scal(mx_adv_ctx* ctx, int lev, const mx_multifab* state,
                          mx_face_multifab* sedge,
                          const mx_face_multifab* umac,
                          const mx_multifab* force, int is_vel,
                          const mx_bcrec* bcs, int nbccomp, int start_scomp,
                          int start_bccomp, int num_comp,
                          int is_conservative);

/* Maestro::MakeRhoXFlux (Maestro.H:230-245, MaestroMakeFlux.cpp:18-465).
   w0 is the Maestro member w0 (edge-centred BaseState). */
int mx_adv_make_rhox_flux(
    mx_adv_ctx* ctx, int lev, mx_face_multifab* sflux,
    mx_multifab* etarhoflux, const mx_face_multifab* sedge,
    const mx_face_multifab* umac, const mx_basestate* rho0_old,
    const mx_basestate* rho0_edge_old, const mx_face_multifab* r0mac_old,
    const mx_basestate* rho0_new, const mx_basestate* rho0_edge_new,
    const mx_face_multifab* r0mac_new,
    const mx_basestate* rho0_predicted_edge, const mx_basestate* w0,
    int start_comp, int num_comp);

/* Maestro::MakeRhoHFlux (Maestro.H:270-295, MaestroMakeFlux.cpp:467-1067). */
int mx_adv_make_rhoh_flux(
    mx_adv_ctx* ctx, int lev, mx_face_multifab* sflux,
    const mx_face_multifab* sedge, const mx_face_multifab* umac,
    const mx_face_multifab* w0mac, const mx_basestate* rho0_old,
    const mx_basestate* rho0_edge_old, const mx_face_multifab* r0mac_old,
    const mx_basestate* rho0_new, const mx_basestate* rho0_edge_new,
    const mx_face_multifab* r0mac_new, const mx_basestate* rhoh0_old,
    const mx_basestate* rhoh0_edge_old, const mx_face_multifab* rh0mac_old,
    const mx_basestate* rhoh0_new, const mx_basestate* rhoh0_edge_new,
    const mx_face_multifab* rh0mac_new, const mx_face_multifab* h0mac_old,
    const mx_face_multifab* h0mac_new);

/* Maestro::CelltoEdge (Maestro.H:401-402, MaestroCelltoEdge.cpp:5-55),
   planar, single radial level, no coarse-fine chunks: edge[r] for
   r = 0..nr on row lev of the BaseStates. */
int mx_adv_cell_to_edge(mx_adv_ctx* ctx, int lev, int nr,
                        const mx_basestate* s0_cell, mx_basestate* s0_edge);

/* Radial geometry of the base state (BaseGeom, Source/BaseGeom.H): the
   radial levels 0..nlev-1 (finest_radial_level + 1), their cell counts
   nr(n) and widths dr(n), and the disjoint chunks [r_start, r_end] each
   level covers (numdisjointchunks, r_start_coord, r_end_coord; chunk i of
   level n is entry i - 1). */
#define MX_MAX_CHUNK 16
typedef struct {
    int nlev;
    int nr_fine;
    int nr[MX_MAX_LEV];
    double dr[MX_MAX_LEV];
    int nchunks[MX_MAX_LEV];
    int r_start[MX_MAX_LEV][MX_MAX_CHUNK];
    int r_end[MX_MAX_LEV][MX_MAX_CHUNK];
} mx_radial_geom;

/* Maestro::MakeEdgeState1d (Maestro.H:1050-1060,
   MaestroMakeEdgeState.cpp:5-981): the time-centred edge states of a
   base-state array s (nlev rows of nr_fine) traced with the base-state
   velocity w0 (the Maestro member, edge-centred) and the source force;
   sedge has rows of nr_fine + 1.  Spherical contexts use row 0 only
   (MakeEdgeState1dSphr); planar ones every level and chunk of rg, with the
   coarse-fine edge synchronisation of MakeEdgeState1dPlanar.  ppm_type,
   slope_order, dt and rel_eps come from the context.  Returns MX_ERR_ARG
   where the reference aborts (a planar ppm_type 1/2 chunk that starts at
   r = 2 or ends at nr - 3). */
int mx_adv_make_edge_state_1d(mx_adv_ctx* ctx, const mx_radial_geom* rg,
                              const mx_basestate* s, mx_basestate* sedge,
                              const mx_basestate* force,
                              const mx_basestate* w0);

/* Maestro::FillUmacGhost (MaestroFillData.cpp:176-453) preceded by the
   periodic FillBoundary of each face MultiFab, as MakeUtrans and Addw0 do
   when finest_level == 0 (MaestroMakeUtrans.cpp:631-640). */
int mx_adv_fill_face_ghosts(mx_adv_ctx* ctx, int lev,
                            mx_face_multifab* uedge);

/* ---- the whole pass ----------------------------------------------------- */

/* One advection pass over one level, the unit of work of the benchmark:
     MakeUtrans -> Addw0(utrans,+1) -> VelPred -> Addw0(umac,+1)
       -> MakeEdgeScal(species) -> MakeEdgeScal(Rho) -> MakeEdgeScal(RhoH)
       -> MakeRhoXFlux -> MakeRhoHFlux
   in the order of AdvancePremac (MaestroAdvection.cpp:7-84),
   DensityAdvance (MaestroDensityAdvance.cpp:115-129,182-184) and
   EnthalpyAdvance (MaestroEnthalpyAdvance.cpp:145-146,229-233), without the
   MAC projection and the producers of the forcing terms, which sit outside
   the path.  scal is the scalar state in predictor form (rho', rhoh', X);
   force terms, base states and outputs are caller-owned.  Old and new base
   states are the same arrays.  Planar runs read the 1-D base states;
   spherical runs read w0mac and the face base states r0mac (rho0), rh0mac
   (rho0 h0) and h0mac (h0) of MakeS0mac (MaestroFill3dData.cpp:464-) in
   MakeRhoXFlux / MakeRhoHFlux. */
typedef struct {
    mx_multifab utilde, ufull, vel_force, w0_cart;
    mx_multifab scal, scal_force;
    mx_face_multifab w0mac;
    mx_face_multifab r0mac, rh0mac, h0mac;
    mx_basestate w0, rho0, rho0_edge, rhoh0, rhoh0_edge, rho0_predicted_edge;
    /* outputs */
    mx_face_multifab utrans, umac, sedge, sflux;
    mx_multifab etarhoflux;
} mx_adv_pass_args;

int mx_adv_advance_pass(mx_adv_ctx* ctx, int lev, const mx_adv_pass_args* a);

/* The same pass with no device allocation.  No counterpart in the reference:
   it is for callers that cannot allocate device memory during the call
   (code under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  Results equal mx_adv_advance_pass on the same arguments.
   The box tables and the temporaries of the staged VelPred kernels live in
   a fixed device array of the library, 256 MiB per GPU (16 MiB of tables,
   240 MiB of workspace), held from the first call on a GPU until the
   process ends.  A call whose tables or workspace do not fit returns
   MX_ERR_SCRATCH before anything is launched.  The fused VelPred kernel (3-D,
   ppm_type 1/2, forces not traced: the default there) needs no workspace;
   the staged kernels need about 192 bytes per cell of the level's boxes
   grown by one (3-D), so levels up to about 1.3 million such cells fit.
   Tables go to the device by stream-ordered copies from host memory the
   context keeps until its next such call, which first waits for the context's
   stream.  Every call on one GPU shares that GPU's array and overwrites it:
   calls that overlap on one GPU (several streams or threads) race.  Must be
   called with the context's device current, and not while a pass graph is
   being created; kernel timing (mx_adv_prof_enable) is not recorded. */
int mx_adv_advance_pass_noalloc(mx_adv_ctx* ctx, int lev, const mx_adv_pass_args* a);

/* ---- a whole level hierarchy (batched) ---------------------------------- */
/* The pass over the hierarchy levels 0..nlev-1 (args[lev] holds level lev's
   fields), with the reference's multi-level semantics (finest_level =
   nlev - 1; reflux_type 1, the default): each stage runs on every level, and
   between the stages the face data are synchronised as the reference does --
   utrans: AverageDownFaces + FillPatchUedge after MakeUtrans and after Addw0
   (MaestroMakeUtrans.cpp:641-647, MaestroFill3dData.cpp:449-455); umac:
   AverageDownFaces after VelPred (MaestroVelPred.cpp:291), AverageDownFaces +
   FillPatchUedge after Addw0; sflux: AverageDownFaces after MakeRhoXFlux and
   MakeRhoHFlux (MaestroMakeFlux.cpp:459-462, 1061-1064).  Cell-centred ghost
   cells of the inputs are filled by the caller, as in the reference.  The
   levels must be properly nested: every ghost face of a level > 0 box lies
   over a coarse face of a coarser-level FAB (its valid faces or its one
   ghost face), as AMReX's grid generation guarantees.  nlev == 1 is
   mx_adv_advance_pass on level 0.  The scalar edge states of all levels go
   to one fused launch per MakeEdgeScal call (species, Rho, RhoH), over one
   box/tile work list. */
int mx_adv_advance_pass_levels(mx_adv_ctx* ctx, int nlev, const mx_adv_pass_args* args);

/* The same hierarchy pass captured once as a HIP graph and replayed with one
   graph launch on the context stream.  Creation runs the pass twice eagerly
   (its outputs are written).  The graph holds the device addresses of the
   caller's buffers as given: replay it only while they stay allocated, and
   destroy it before the context. */
typedef struct mx_adv_graph mx_adv_graph;
int mx_adv_pass_graph_create(mx_adv_ctx* ctx, int nlev, const mx_adv_pass_args* args,
                             mx_adv_graph** out);
int mx_adv_pass_graph_launch(mx_adv_graph* g);
int mx_adv_pass_graph_destroy(mx_adv_graph* g);

/* Autotuner of the fused edge-state kernel (no counterpart in the
   reference): runs the hierarchy pass once, then times the species edge-state
   launch for every tile shape (16x11, 16x12, 16x8, 32x8 columns) and z-chunk (16,
   32, 64, 128 planes), keeps the fastest in the context ("edge_scal_tile",
   "edge_scal_tz") and reports every candidate. */
#define MX_TUNE_MAX 16
typedef struct {
    int ncand;
    int tile[MX_TUNE_MAX];
    int tz[MX_TUNE_MAX];
    double ms[MX_TUNE_MAX];
    int best; /* index of the chosen candidate */
} mx_adv_tune_report;
int mx_adv_autotune(mx_adv_ctx* ctx, int nlev, const mx_adv_pass_args* args,
                    mx_adv_tune_report* report);

/* ---- base state <-> Cartesian producers --------------------------------- */
/* The spherical geometry these read (Maestro members center, geom.ProbLo,
   base_geom and the *_interp_type runtime parameters).  r_cc_loc and
   r_edge_loc are row 0 of base_geom.r_cc_loc / r_edge_loc (nr_fine and
   nr_fine + 1 values) in memory the library's kernels can read (device
   memory).  Only the non-exact base state (use_exact_base_state = 0, radial
   index int(radius / dr_fine)) is supported: the exact one maps cells to
   radii through the cell_cc_to_r integer MultiFab. */
typedef struct {
    double center[3];
    double prob_lo[3];
    double dr_fine;
    int nr_fine;
    const double* r_cc_loc;
    const double* r_edge_loc;
    int s0_interp_type;    /* 1-3 (piecewise constant, linear, quadratic) */
    int w0_interp_type;    /* 1-3 */
    int s0mac_interp_type; /* 1-3 */
    int w0mac_interp_type; /* 1-4 */
} mx_sph_geom;

/* Maestro::Put1dArrayOnCart, one level (Maestro.H, MaestroFill3dData.cpp:49-339):
   the base state s0 (bin- or edge-centred) put on the valid cells of s0_cart:
   component dim - 1 of a vector or component 0 of a scalar (planar), or all
   three components of a vector along the radius (spherical, sg required).
   The reference's multi-level wrapper (:9-38) then calls AverageDown and
   FillPatch; those stay with the caller. */
int mx_adv_put_1d_array_on_cart(mx_adv_ctx* ctx, int lev, const mx_basestate* s0,
                                mx_multifab* s0_cart, int is_input_edge_centered,
                                int is_output_a_vector, const mx_sph_geom* sg);

/* Maestro::MakeW0mac, one level (MaestroFill3dData.cpp:464-710), spherical:
   w0mac on every face of each box grown by one cell (w0mac needs 1 ghost).
   w0mac_interp_type 1 averages w0_cart (3 components, 2 ghost cells filled
   by the caller); 2 and 3 interpolate the edge-centred w0 to the face; 4
   averages w0 interpolated to the face's four nodes. */
int mx_adv_make_w0mac(mx_adv_ctx* ctx, int lev, const mx_basestate* w0,
                      const mx_multifab* w0_cart, mx_face_multifab* w0mac,
                      const mx_sph_geom* sg);

/* Maestro::MakeS0mac, one level (MaestroFill3dData.cpp:712-1188), spherical:
   s0mac on every face of each box grown by one cell.  s0mac_interp_type 1
   averages s0_cart, which the reference builds with Put1dArrayOnCart and
   fills with FillPatch: the caller passes it (1 component, 2 ghost cells);
   2 and 3 interpolate the bin-centred s0 to the face (s0_cart unused). */
int mx_adv_make_s0mac(mx_adv_ctx* ctx, int lev, const mx_basestate* s0,
                      const mx_multifab* s0_cart, mx_face_multifab* s0mac,
                      const mx_sph_geom* sg);

/* Maestro::PutInPertForm, one level (MaestroConvert.cpp:36-62): scal[comp]
   -= s0 on the cells (flag != 0) or += s0 (flag == 0), valid cells; s0 is
   put on the cells as Put1dArrayOnCart does (bin-centred, scalar).  The
   reference's average_down and FillPatch stay with the caller. */
int mx_adv_put_in_pert_form(mx_adv_ctx* ctx, int lev, mx_multifab* scal,
                            const mx_basestate* s0, int comp, int flag,
                            const mx_sph_geom* sg);

/* Maestro::MakeVelForce, one level (Maestro.H, MaestroForce.cpp:6-258),
   without ROTATION (a compile-time option of the reference, off by default):
   vel_force on the valid cells from the buoyancy of rho - rho0, -gpi / rho,
   the w0 force and, with do_add_utilde_force, the Utilde . grad w0 term.
   rho is component 0 of the given MultiFab (the state); gpi,
   w0_force_cart: dim components; w0_cart: dim components with 1 ghost cell
   (planar, read with do_add_utilde_force); normal: 3 components
   (spherical).  rho0 and grav_cell are put on the cells as Put1dArrayOnCart
   does; gradw0 = (w0(r+1) - w0(r)) / dr_fine is zero with
   average_base_state or the exact base state, as in the reference.
   buoyancy_cutoff_factor, base_cutoff_density, do_add_utilde_force and
   average_base_state are Maestro members there; do_add_utilde_force must be
   0 or 1 (the reference reads it as a truth value in the planar branch and as
   == 1 in the spherical one).  The ghosts of vel_force
   (zeroed, then FillPatch in the reference) stay with the caller. */
int mx_adv_make_vel_force(mx_adv_ctx* ctx, int lev, mx_multifab* vel_force,
                          const mx_face_multifab* uedge, const mx_multifab* rho,
                          const mx_multifab* gpi, const mx_multifab* w0_force_cart,
                          const mx_multifab* w0_cart, const mx_multifab* normal,
                          const mx_basestate* rho0, const mx_basestate* grav_cell,
                          const mx_basestate* w0, double buoyancy_cutoff_factor,
                          double base_cutoff_density, int do_add_utilde_force,
                          int average_base_state, const mx_sph_geom* sg);

/* Maestro::ModifyScalForce, one level (MaestroForce.cpp:261-466): component
   comp of scal_force gains -s div(U) (fullform) or -(s - s0) div(U) -
   div(s0 U) on the valid cells, with div(U) including w0 (planar: w0_cart,
   dim components, 1 ghost cell; spherical: divw0 of w0 put on the cells).
   s0_cart: 1 component (spherical, not fullform: 1 ghost cell).  Planar
   s0_edge is put on the cells as the edge below each cell; the cell above a
   box's top cell takes that value inside the domain and, past the domain
   top, the top cell's value (non-periodic: the first-order extrapolation of
   the reference's FillPatch at level 0) or the periodic image. */
int mx_adv_modify_scal_force(mx_adv_ctx* ctx, int lev, mx_multifab* scal_force,
                             const mx_multifab* state, const mx_face_multifab* umac,
                             const mx_basestate* s0_edge, const mx_multifab* s0_cart,
                             const mx_multifab* w0_cart, const mx_basestate* w0, int comp,
                             int fullform, const mx_sph_geom* sg);

/* Maestro::ConvertRhoXToX (MaestroConvert.cpp:64-91) and ConvertRhoHToH
   (:93-114), one level: the species (or RhoH) divided by Rho (flag != 0) or
   multiplied by it (flag == 0) on the valid cells.  AverageDown and
   FillPatch stay with the caller. */
int mx_adv_convert_rhox_to_x(mx_adv_ctx* ctx, int lev, mx_multifab* scal, int flag);
int mx_adv_convert_rhoh_to_h(mx_adv_ctx* ctx, int lev, mx_multifab* scal, int flag);

/* ---- Ghost fill ---------------------------------------------------------- */
/* Maestro::FillPatch(lev, time, mf, mf_old, mf_new, srccomp, destcomp, ncomp,
   startbccomp, bcs_in, variable_type) at level 0 (MaestroFillData.cpp:24-47):
   amrex::FillPatchSingleLevel with ScalarFill / VelFill and FillExtBC
   (MaestroBCFill.cpp:8-171).  `src` is the MultiFab GetData picks for `time`
   (mf_new or mf_old; it may be `mf` itself); `bcs` points at the ncomp
   records bcs_in[startbccomp ..]; variable_type 1 is velocity.
     1. every cell of mf's FABs (valid and ghost) that lies in a valid box of
        src, or in a periodic image of one, takes src's value (srccomp.. ->
        destcomp..);
     2. cells beyond a domain side take fab_filcc's rule for their
        component's BCRec (foextrap, hoextrap, reflect_even, reflect_odd;
        int_dir and ext_dir leave them), direction x, then y, then z, for
        every FAB whose grown box leaves the domain grown in the periodic
        directions;
     3. beyond an ext_dir side Pi and velocity become 0; any other scalar
        with an ext_dir side returns MX_ERR_BC (the reference aborts) before
        anything is written.
   Every FAB of mf is its valid box grown by one ghost width.  Level > 0
   is mx_adv_fill_patch_two_levels; here it returns MX_ERR_UNSUPPORTED.
   There is no time argument: a time strictly between t_old and t_new, for
   which GetData hands over both MultiFabs, has no counterpart here. */
int mx_adv_fill_patch(mx_adv_ctx* ctx, int lev, mx_multifab* mf, const mx_multifab* src,
                      int srccomp, int destcomp, int ncomp, const mx_bcrec* bcs,
                      int variable_type);

/* Maestro::FillPatch at a level lev > 0 (MaestroFillData.cpp:48-66):
   amrex::FillPatchTwoLevels with cell_cons_interp and the boundary functions
   of mx_adv_fill_patch.  `crse` is level lev - 1's and `fine` level lev's
   MultiFab that GetData picks (`fine` may be `mf`; `crse` needs its valid
   cells only).  Steps:
     1. fine patches: per FAB of mf, its cells in the fine domain (grown
        without bound in the periodic directions) that no fine valid box or
        periodic image holds;
     2. per fine patch a coarse patch, the patch coarsened and grown by one
        coarse cell, filled as FillPatchSingleLevel on level lev - 1 does
        (copy from crse's valid boxes and images; the boundary rules with
        the coarse domain, the patch's components numbered 0..ncomp-1 for
        FillExtBC's Pi test, as in the reference);
     3. CellConservativeLinear with linear limiting: per coarse cell and
        direction the central slope (one-sided at a domain side with
        ext_dir / hoextrap the coarse patch reaches), bounded by twice the
        one-sided differences (zero at an extremum), and one factor per
        direction shared by the ncomp components, the minimum of 1 and the
        bounded-over-central ratios; a fine cell gets the coarse value plus
        (i - ic r + 1/2) / r - 1/2 times each limited slope, x, y, z in order;
     4. FillPatchSingleLevel on level lev from `fine` (mx_adv_fill_patch's
        steps).
   A coarse patch cell inside the domain that no level lev - 1 box holds
   (levels not properly nested for this ghost width) returns MX_ERR_BOX
   before anything is written; so do the checks of mx_adv_fill_patch.
   ncomp is at most 32. */
int mx_adv_fill_patch_two_levels(mx_adv_ctx* ctx, int lev, mx_multifab* mf,
                                 const mx_multifab* crse, const mx_multifab* fine, int srccomp,
                                 int destcomp, int ncomp, const mx_bcrec* bcs,
                                 int variable_type);

/* ---- Plotfile -------------------------------------------------------------- */
/* amrex::WriteMultiLevelPlotfile(plotfilename, nlevels, mf, varnames, geom,
   time, level_steps, ref_ratio) as Maestro::WritePlotFile calls it
   (MaestroPlot.cpp:99-116): the AMReX plotfile directory `dir` (format
   HyperCLaw-V1.1; Header, Level_<l>/Cell_H, Level_<l>/Cell_D_00000) holding
   components 0..ncomp-1 of the valid cells of mf[0..nlev-1].  Geometry and
   boxes are the context's; prob_lo/prob_hi are geom[0].ProbLo()/ProbHi().
   Cartesian coordinates, one data file per level, little-endian doubles.
   Returns MX_ERR_IO when a file cannot be written.  Host code: it waits for
   the context stream. */
int mx_adv_write_plotfile(mx_adv_ctx* ctx, const char* dir, int nlev,
                          const mx_multifab* const* mf, int ncomp, const char* const* varnames,
                          double time, const int* level_steps, const double* prob_lo,
                          const double* prob_hi);

/* ---- AMR restriction ---------------------------------------------------- */
/* amrex::average_down(fine, crse, ..., comp, ncomp, ratio) between levels
   lev + 1 (fine) and lev (crse), the step of Maestro::AverageDown
   (MaestroFillData.cpp:145-154, which loops lev = finest_level - 1 .. 0):
   every coarse valid cell covered by the coarsened fine BoxArray becomes
   the mean of its r^dim fine cells (summed k outermost, i innermost, times
   1 / r^dim); other coarse cells and the coarse ghosts keep their values.
   The ratio is the ratio of the two levels' domains. */
int mx_adv_average_down(mx_adv_ctx* ctx, int lev, const mx_multifab* fine, mx_multifab* crse,
                        int comp, int ncomp);

/* amrex::average_down_faces(fine, crse, ratio) between levels lev + 1 and
   lev, the step of Maestro::AverageDownFaces (MaestroFillData.cpp:156-172):
   every coarse face covered by the coarsened fine faces becomes the mean of
   the r^(dim-1) fine faces on it, every component.  A face shared by two
   fine boxes must hold one value in both (as after FillBoundary). */
int mx_adv_average_down_faces(mx_adv_ctx* ctx, int lev, const mx_face_multifab* fine,
                              mx_face_multifab* crse);

/* Maestro::FillPatchUedge (MaestroFillData.cpp:455-833) over levels
   0..nlev-1 (uedge[lev] holds level lev's one-component face data, one
   ghost face): on every level lev > 0, the ghost faces no box of the level
   holds are filled from level lev - 1 (piecewise constant on the fine faces
   of each coarse face, the level's own valid faces where they exist, and
   linear interpolation in the normal direction between coarse-aligned
   faces, piecewise constant across); then, on every level, FillBoundary
   with periodicity and FillUmacGhost.  Coarse faces no coarse FAB holds read
   as 1e200, as in the reference. */
int mx_adv_fill_patch_uedge(mx_adv_ctx* ctx, int nlev, mx_face_multifab* uedge);

/* ---- flux registers (reflux_type 2) -------------------------------------- */
/* Maestro::flux_reg_s[lev] (Maestro.H:2021-2025), an amrex::FluxRegister on
   level lev's BoxArray for the lev / lev - 1 interface.  The context holds
   one register per level lev >= 1; mx_adv_set_level defines it with Nscal
   components when params.reflux_type == 2 (MaestroInit.cpp:389-392), or
   mx_adv_flux_reg_define does so explicitly (the FluxRegister constructor).
   The register holds, per fine box and side, the faces of the coarsened box
   on that side; values start at 0.

   With reflux_type == 2, mx_adv_make_rhox_flux / mx_adv_make_rhoh_flux and
   the passes update the registers as MakeRhoXFlux / MakeRhoHFlux do
   (MaestroMakeFlux.cpp:420-451, 1035-1058): on level lev, CrseInit into the
   register of lev + 1 with mult = -dt * area[d] and FineAdd into the
   register of lev with mult = dt * area[d], area the level's face areas;
   components start_comp.. and Rho (MakeRhoXFlux) or RhoH (MakeRhoHFlux).
   The caller then calls mx_adv_flux_reg_reflux where the reference's
   UpdateScal calls Reflux (MaestroAdvection.cpp:258-278). */
int mx_adv_flux_reg_define(mx_adv_ctx* ctx, int lev, int ncomp);

/* FluxRegister::setVal: every register value of level lev becomes v. */
int mx_adv_flux_reg_setval(mx_adv_ctx* ctx, int lev, double v);

/* FluxRegister::CrseInit(mflx, dir, srccomp, destcomp, numcomp, mult) with
   mflx = level lev - 1's fluxes of direction dir (one mx_multifab, a FAB per
   coarse box holding its valid faces): every register face of direction dir
   becomes mflx * mult, taken from the first coarse box (box order) whose
   valid faces hold it. */
int mx_adv_flux_reg_crse_init(mx_adv_ctx* ctx, int lev, const mx_multifab* flux, int dir,
                              int srccomp, int destcomp, int numcomp, double mult);

/* FluxRegister::FineAdd(mflx, dir, srccomp, destcomp, numcomp, mult) with
   mflx = level lev's fluxes of direction dir: every register face of a fine
   box adds mult * flux of each fine face on it, one addition per fine face,
   fine face loops (k, j) for dir 0, (k, i) for dir 1, (j, i) for dir 2,
   the first outermost. */
int mx_adv_flux_reg_fine_add(mx_adv_ctx* ctx, int lev, const mx_multifab* flux, int dir,
                             int srccomp, int destcomp, int numcomp, double mult);

/* FluxRegister::Reflux(state, scale, scomp, dcomp, ncomp, geom[lev - 1]) on
   level lev - 1's state: for every valid coarse cell and each orientation in
   order (the low sides of x, y, z, then the high sides), f = the register
   value on the cell's face across that orientation (its high face for a low
   side, its low face for a high side; periodic images included; 0 where no
   register face lies) and s += (-scale * f) / vol on a low side,
   s += (scale * f) / vol on a high side, vol = the coarse cell volume. */
int mx_adv_flux_reg_reflux(mx_adv_ctx* ctx, int lev, mx_multifab* state, double scale,
                           int scomp, int dcomp, int ncomp);

/* Copy every register value of level lev to host memory (test access; no
   counterpart in the reference): registers in (fine box, orientation) order,
   each with its components outermost and faces i fastest.  *n receives the
   count; host may be NULL to query it. */
int mx_adv_flux_reg_read(mx_adv_ctx* ctx, int lev, double* host, long long cap, long long* n);

/* ---- the update that closes the step ------------------------------------ */
/* Maestro::UpdateScal (Maestro.H:306-311, MaestroAdvection.cpp:86-297) on
   level lev: statenew = stateold + dt (-div sflux + force) on the valid
   cells, for start_comp == RhoH (num_comp 1) or FirstSpec (num_comp NumSpec;
   then also Rho as the sum of the species updates, the auxiliary components
   FirstSpec + NumSpec .. nscal - 1 rescaled by the density, the density floor
   0.5 base_cutoff_density and the redistribution of negative species).  The
   reference reads base_cutoff_density and do_eos_h_above_cutoff as members;
   here they are arguments.  do_eos_h_above_cutoff != 0 returns
   MX_ERR_UNSUPPORTED: its branch calls the equation of state, which is not
   part of this library; p0_cart is read only by that branch (NULL is fine).
   The reference's reflux, AverageDown and FillPatch after the update stay
   with the caller.  All MultiFabs hold nscal components; no ghosts needed. */
int mx_adv_update_scal(mx_adv_ctx* ctx, int lev, const mx_multifab* stateold,
                       mx_multifab* statenew, const mx_face_multifab* sflux,
                       const mx_multifab* force, int start_comp, int num_comp,
                       const mx_multifab* p0_cart, double base_cutoff_density,
                       int do_eos_h_above_cutoff);

/* Maestro::UpdateVel (Maestro.H:320-326, MaestroAdvection.cpp:299-526) on
   level lev: unew = uold - dt (Utilde . grad) Utilde + dt force - the
   (w0 . grad) Utilde term, times the sponge when do_sponge != 0, on the valid
   cells.  uold and unew are the Maestro members of the same names; w0_cart
   (planar: dim components, 1 ghost cell, read in the cell above each valid
   cell as the reference does) and do_sponge are members there too.  w0mac
   is read in spherical contexts only, w0_cart in planar ones only.
   AverageDown and FillPatch stay with the caller. */
int mx_adv_update_vel(mx_adv_ctx* ctx, int lev, const mx_multifab* uold, mx_multifab* unew,
                      const mx_face_multifab* umac, const mx_face_multifab* uedge,
                      const mx_multifab* force, const mx_multifab* sponge,
                      const mx_face_multifab* w0mac, const mx_multifab* w0_cart, int do_sponge);

/* ---- differentiation ---------------------------------------------------- */
/* A differentiable scalar advection step (no single counterpart: the
   enthalpy step of EnthalpyAdvance with enthalpy_pred_type predict_rhohprime,
   the default, for a perturbation (rho h)' with a zero base state, a fixed
   MAC velocity and no forcing: MakeEdgeScal in the advective form
   (MaestroEnthalpyAdvance.cpp:143-147), MakeRhoHFlux's
   umac (rhoh0_edge + sedge) with rhoh0_edge = 0 (MaestroMakeFlux.cpp:699-708)
   and UpdateScal's enthalpy update before its EOS branch
   (MaestroAdvection.cpp:120-135)):
     sedge = MakeEdgeScal(s, umac, force 0, advective form), ppm_type 1 or 2
     snew  = s + dt (-div(umac sedge) + 0)
   A non-zero base state adds a flux that does not depend on s, so the
   state adjoint is the same; the velocity adjoint leaves that flux out.
   is_conservative != 0 predicts in the conservative form instead
   (MaestroMakeEdgeScalars.cpp:906-929, 1411-1438, MakeDivU 648-660): the
   species step with species_pred_type predict_rhoX
   (MaestroDensityAdvance.cpp:118-122, MaestroMakeFlux.cpp:134-137) and the
   enthalpy step with predict_rhoh (MaestroEnthalpyAdvance.cpp:139-142).
   on a context whose level 0 is one fully periodic box [0, n)^3.  s holds
   ncomp components with >= 3 ghost cells (>= 4 for ppm_type 2; fewer is
   MX_ERR_BOX), umac[0..2] one face FAB per
   direction with >= 1 ghost face.  The call fills the ghost cells of s and
   umac by periodic copy and sets face n of umac to face 0 (one periodic face
   carries one value). */
int mx_adv_scal_step(mx_adv_ctx* ctx, mx_fab* s, mx_fab* umac, mx_fab* snew, int ncomp,
                     int is_conservative);

/* Hand-written adjoint of mx_adv_scal_step (no counterpart in the
   reference): given gbar = dL/dsnew on the valid cells, writes sbar = dL/ds
   on the valid cells and ubar[d] = dL/dumac[d] on faces 0..n-1 (face n, if
   the FAB holds it, is set to 0: it is face 0).  It reads only the valid
   cells of s and faces 0..n-1 of umac (no ghosts needed); inputs that are
   dense (ncomp, n, n, n) arrays starting at index 0 are read in place.
   Piecewise differentiation: every limiter and upwind branch is taken as in
   the forward step.  Deterministic (no atomics); the workspace holds about
   65 ncomp + 3 arrays of n^3 doubles. */
int mx_adv_scal_step_vjp(mx_adv_ctx* ctx, mx_fab* s, mx_fab* umac, const mx_fab* gbar,
                         mx_fab* sbar, mx_fab* ubar, int ncomp, int is_conservative);

/* The same step over the level-0 box list of a context (no single
   counterpart: the composition above on MultiFabs, as DensityAdvance runs
   it per FAB).  The boxes must tile the domain; each direction is periodic
   or has outflow on both sides (phys_bc Outflow: foextrap ghost cells and
   the foextrap rules at the domain faces, MaestroMakeEdgeScalars.cpp:
   709-738, 931-961, 1446-1480); other boundaries return
   MX_ERR_UNSUPPORTED.  s needs >= 3 ghost cells (>= 4 for ppm_type 2),
   umac >= 1 ghost face; the call fills them (FillPatch at level 0, then
   FillBoundary + FillUmacGhost).  A face two boxes share must hold one
   value in both. */
int mx_adv_scal_step_boxes(mx_adv_ctx* ctx, mx_multifab* s, mx_face_multifab* umac,
                           mx_multifab* snew, int ncomp, int is_conservative);

/* Hand-written adjoint of mx_adv_scal_step_boxes (no counterpart in the
   reference): gbar = dL/dsnew on the valid cells gives sbar = dL/ds on the
   valid cells and ubar[d] = dL/dumac[d] on every face of direction d a box
   holds (faces shared by boxes, and periodic images, all get the one
   value of their face).  It reads the valid cells of s and the faces
   lo..hi of umac (and the face at an outflow high side); no ghosts needed.
   The state adjoint of a cell on an outflow side includes the ghost cells
   that copy it.  Piecewise differentiation as in mx_adv_scal_step_vjp.
   Deterministic.  Workspace: about 50 ncomp + 8 (ppm_type 2: 62 ncomp + 8)
   arrays of the domain grown by 7 points per outflow direction. */
int mx_adv_scal_step_vjp_boxes(mx_adv_ctx* ctx, const mx_multifab* s,
                               const mx_face_multifab* umac, const mx_multifab* gbar,
                               mx_multifab* sbar, mx_face_multifab* ubar, int ncomp,
                               int is_conservative);

#ifdef __cplusplus
}
#endif

#endif /* MX_ADV_H */
