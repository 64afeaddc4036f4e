// mx_internal.H -- host-side context, workspace and launch helpers of
// libmaestroex_amd.so (not part of the C-ABI).
#ifndef MX_INTERNAL_H
#define MX_INTERNAL_H

#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "mx_adv.h"
#include "mx_device.H"

struct mx_adv_ctx {
    mx_adv_params p;
    std::vector<mx_box> grids[MX_MAX_LEV];
    std::vector<mx_bcrec> bcs_s, bcs_u;
    hipStream_t stream = nullptr;
    int device = 0;

    // device workspace: a bump allocator reset at the start of every entry
    // point (all work is ordered on `stream`, so a later call reuses memory
    // only after the earlier kernels have finished with it)
    char* ws = nullptr;
    size_t ws_cap = 0;
    size_t ws_off = 0;

    // the box-list adjoint's padded arrays: no kernel writes their padding,
    // so it stays zero from one call to the next while the layout
    // (adj_layout) is unchanged
    char* adj_buf = nullptr;
    size_t adj_cap = 0;
    long long adj_layout[8] = {0, 0, 0, 0, 0, 0, 0, 0};

    // per-call box tables: built in htab; a new table is copied through a
    // ring of pinned staging buffers, so that an upload never waits for the
    // stream to drain (a pageable copy would).  A slot is reused after its
    // previous copy has completed (its event).
    struct TabSlot {
        char* host = nullptr;
        size_t cap = 0;
        hipEvent_t done = nullptr;
        bool pending = false;
    };
    static constexpr int kTabSlots = 32;
    std::vector<TabSlot> tab_slots;
    int tab_next = 0;
    std::vector<char> htab;
    // device copies of the tables already uploaded, found by content: a pass
    // over unchanged grids and arrays builds the same tables every call and
    // then copies nothing.  Entries are never written after their upload.
    struct TabEntry {
        unsigned long long hash;
        std::vector<char> bytes;
        char* dev;
    };
    std::vector<TabEntry> tab_cache;
    size_t tab_cache_bytes = 0;

    // FillBoundary plans, keyed by (lev, nodal dir, ghost width): the ghost
    // shell of every box as slabs, each with the box's candidate sources
    // (source box, periodic shift) in the order that decides shared faces
    struct GhostSlab {
        int box;
        int lo[3], hi[3];
        int cand_off, cand_n;
    };
    struct GhostCand {
        int src_box;
        int shift[3];
    };
    struct GhostPlan {
        std::vector<GhostSlab> slabs;
        std::vector<GhostCand> cands;
    };
    std::map<std::tuple<int, int, long long>, GhostPlan> plans;
    // FillPatch copy plans, keyed by (lev, ghost widths): every intersection
    // of a destination FAB's grown box with a source valid box or one of its
    // periodic images
    struct FillCopy {
        int b, f;
        int sh[3];
        int lo[3], hi[3];
    };
    std::map<std::tuple<int, int, int, int>, std::vector<FillCopy>> fill_plans;

    // flux registers (amrex::FluxRegister) of levels 1..finest_level: per
    // fine box b and orientation o (the low sides of the directions, then the
    // high sides) the faces of the coarsened box on that side, in coarse face
    // index space; box[b * 2 * dim + o], values at d + off[...] (components
    // outermost, i fastest)
    struct FluxRegDev {
        int ncomp = 0;
        std::vector<mx_box> box;
        std::vector<long long> off;
        long long total = 0;
        double* d = nullptr;
    };
    FluxRegDev freg[MX_MAX_LEV];

    // opt-in kernel timing: a HIP event pair around every launch
    struct ProfRec {
        const char* name;
        hipEvent_t a, b;
    };
    int prof_on = 0;
    std::vector<ProfRec> prof_recs;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::string> prof_names;
    std::vector<double> prof_ms;
    std::vector<long long> prof_n;

    // kernel selection (mx_adv_set_tuning)
    int fused_edge = 1;
    int fused_tz = 32;
    int vp_fused_xcd = -1; // fused VelPred: XCD-aware block order (xcd_tile mode); -1: 2 for
                           // one level of equal boxes, else 1
    int es_fused_xcd = 1;  // fused MakeEdgeScal: XCD-aware block order (xcd_tile mode)
    int xcd_remap = 0;  // KGeom::xcd of the staged pass kernels
    int vp_ppm_all = 1;  // VelPred's PPM kernel: 1 = all components per thread
    int utrans_line = 1;   // MakeUtrans (3-D, ppm_type 1/2): 1 = each cell traced once
    int utrans_chunks = 8; // chunks per y/z column of that form
    int vp_trans_all = 1;  // VelPred's transverse kernel: 1 = all six states per thread
    int vp_fused = 1;      // VelPred: 1 = the fused z-marching kernel where it applies
    int vp_fused_tz = 32;  // planes per chunk of the fused VelPred kernel
    int vp_fused_comp = 1; // fused VelPred: 1 = one velocity component per block
    int rhoh_flux_epilogue = 1;  // the pass: MakeRhoHFlux (planar, ept 0/1/3) in the
                                 // fused edge kernel's epilogue
    int vp_fused_tile = 2; // its tile (waves per SIMD): 0 16x11 (3), 1 16x8 (3), 2 16x11 (4), 3 32x8 (3)
    int fused_tile = -1;    // 0: 32 x 8 columns, 1: 16 x 16, 2: 16 x 12, 3: 16 x 8,
                            // 4: 16 x 11; -1: 16 x 11 for ppm_type 1, else 16 x 12
    int fused_per_box = 0;  // 1: one fused launch per box (comparison mode)
    int adj_state_fused = 1;  // adjoint: face, slope and state adjoints in one gather
    int adj_face_sums = 0;    // adjoint: predictor-state face adjoints stored once per face
    int adj_pow2 = 1;         // adjoint: mask indexing when n is a power of two
    int adj_trace_rolled = 0; // adjoint: trace kernel with one direction per loop trip
    int adj_xcd = 0;          // adjoint: each XCD takes a contiguous range of blocks

    // calls that allocate nothing (mx_adv_advance_pass_noalloc).  With `dry`
    // set an entry point runs its host logic only: every launch is skipped,
    // a table upload adds its size to dry_tab and a workspace request raises
    // dry_ws.  With ext_tab / ext_ws set, tables and workspace are placed in
    // that caller-given device memory instead of the context's own; the host
    // bytes of each table stay in ext_host until the stream has passed the
    // copies (the next such call synchronises first).
    int dry = 0;
    size_t dry_tab = 0, dry_ws = 0;
    char* ext_tab = nullptr;
    size_t ext_tab_cap = 0, ext_tab_off = 0;
    char* ext_ws = nullptr;
    size_t ext_ws_cap = 0;
    std::vector<std::vector<char>> ext_host;

    // box tables under graph capture: 0 eager (one device table, copied at
    // every call); 1 record (eager, and keep each table's bytes in order);
    // 2 replay (each table is the next recorded one, already in rec_arena)
    int tab_mode = 0;
    std::vector<std::vector<char>> rec_tabs;
    std::vector<size_t> rec_offs;
    char* rec_arena = nullptr;
    size_t rec_next = 0;
};

namespace mxh {

#define MX_HIP_CHECK(x)                                  \
    do {                                                 \
        hipError_t e_ = (x);                             \
        if (e_ != hipSuccess) return MX_ERR_HIP;         \
    } while (0)

hipEvent_t prof_event(mx_adv_ctx* c);

// launch on the context stream; with timing on, bracket it with events
#define MX_LAUNCH(c, name, kern, grid, block, ...)                                  \
    do {                                                                            \
        if ((c)->dry) break;                                                        \
        hipEvent_t e0_ = nullptr;                                                   \
        if ((c)->prof_on) {                                                         \
            e0_ = mxh::prof_event(c);                                               \
            MX_HIP_CHECK(hipEventRecord(e0_, (c)->stream));                         \
        }                                                                           \
        hipLaunchKernelGGL(kern, grid, block, 0, (c)->stream, __VA_ARGS__);         \
        MX_HIP_CHECK(hipGetLastError());                                            \
        if ((c)->prof_on) {                                                         \
            hipEvent_t e1_ = mxh::prof_event(c);                                    \
            MX_HIP_CHECK(hipEventRecord(e1_, (c)->stream));                         \
            (c)->prof_recs.push_back({name, e0_, e1_});                             \
        }                                                                           \
    } while (0)

inline mxd::DA to_da(const mx_fab& f) {
    mxd::DA a;
    a.p = f.p;
    a.js = f.jstride;
    a.ks = f.kstride;
    a.ns = f.nstride;
    a.b0 = f.begin[0];
    a.b1 = f.begin[1];
    a.b2 = f.begin[2];
    return a;
}

inline mxd::DA null_da() {
    mxd::DA a;
    a.p = nullptr;
    a.js = a.ks = a.ns = 0;
    a.b0 = a.b1 = a.b2 = 0;
    return a;
}

struct HBox {
    int lo[3], hi[3];
};

inline HBox hbox(const mx_box& b) {
    HBox h;
    for (int d = 0; d < 3; ++d) {
        h.lo[d] = b.lo[d];
        h.hi[d] = b.hi[d];
    }
    return h;
}
inline HBox grow(HBox b, int dim, int n) {
    for (int d = 0; d < dim; ++d) {
        b.lo[d] -= n;
        b.hi[d] += n;
    }
    return b;
}
inline HBox grow_dir(HBox b, int d, int n) {
    b.lo[d] -= n;
    b.hi[d] += n;
    return b;
}
inline HBox grow_hi(HBox b, int d, int n) {
    b.hi[d] += n;
    return b;
}
inline HBox grow_lo(HBox b, int d, int n) {
    b.lo[d] -= n;
    return b;
}
inline HBox nodal(HBox b, int d) {
    b.hi[d] += 1;
    return b;
}
inline long long npts(const HBox& b) {
    long long n = 1;
    for (int d = 0; d < 3; ++d) n *= (long long)(b.hi[d] - b.lo[d] + 1);
    return n;
}

// does FAB f cover region r for components [0, ncomp)?
inline bool covers(const mx_fab& f, const HBox& r, int ncomp) {
    if (!f.p) return false;
    if (f.ncomp < ncomp) return false;
    for (int d = 0; d < 3; ++d) {
        if (r.lo[d] < f.begin[d] || r.hi[d] >= f.end[d]) return false;
    }
    return true;
}

mxd::KGeom kgeom(const mx_adv_ctx* c, int lev);

// workspace
int ws_reserve(mx_adv_ctx* c, size_t bytes);
void free_tab_slot(mx_adv_ctx::TabSlot& t);
void free_tab_cache(mx_adv_ctx* c);
double* ws_take(mx_adv_ctx* c, size_t ndouble);

// a temporary FAB per box over `region(box)` with ncomp components
template <class RegionFn>
std::vector<mxd::DA> tmp_fabs(mx_adv_ctx* c, const std::vector<mx_box>& grids, int ncomp,
                              RegionFn region) {
    std::vector<mxd::DA> out;
    for (const mx_box& g : grids) {
        HBox r = region(hbox(g));
        long long n0 = r.hi[0] - r.lo[0] + 1, n1 = r.hi[1] - r.lo[1] + 1,
                  n2 = r.hi[2] - r.lo[2] + 1;
        mxd::DA a;
        a.p = ws_take(c, (size_t)(n0 * n1 * n2 * ncomp));
        a.js = n0;
        a.ks = n0 * n1;
        a.ns = n0 * n1 * n2;
        a.b0 = r.lo[0];
        a.b1 = r.lo[1];
        a.b2 = r.lo[2];
        out.push_back(a);
    }
    return out;
}

template <class RegionFn>
size_t tmp_bytes(const std::vector<mx_box>& grids, int ncomp, RegionFn region) {
    size_t s = 0;
    for (const mx_box& g : grids) {
        HBox r = region(hbox(g));
        s += (size_t)npts(r) * ncomp * sizeof(double) + 256;
    }
    return s;
}

// per-call device table: append host structs, upload once, get device pointers
struct Table {
    mx_adv_ctx* c;
    char* base = nullptr;  // device copy, set by upload()
    explicit Table(mx_adv_ctx* ctx) : c(ctx) { c->htab.clear(); }
    template <class T>
    size_t add(const std::vector<T>& v) {
        size_t off = (c->htab.size() + 255) & ~(size_t)255;
        c->htab.resize(off + v.size() * sizeof(T));
        std::memcpy(c->htab.data() + off, v.data(), v.size() * sizeof(T));
        return off;
    }
    int upload();
    template <class T>
    const T* dev(size_t off) const {
        return reinterpret_cast<const T*>(base + off);
    }
};

// launch geometry: x covers the largest per-box region, y = boxes, z = extra
inline dim3 grid_for(long long maxpts, int nbox, int nz, int block) {
    long long gx = (maxpts + block - 1) / block;
    if (gx < 1) gx = 1;
    return dim3((unsigned)gx, (unsigned)nbox, (unsigned)nz);
}

// grid_for for kernels that take their point index from grid_point: with
// the XCD remap on, x is rounded up to a multiple of 8 (the extra blocks'
// points lie past the region and return)
inline dim3 grid_for_g(const mxd::KGeom& g, long long maxpts, int nbox, int nz, int block) {
    dim3 d = grid_for(maxpts, nbox, nz, block);
    if (g.xcd) d.x = (d.x + 7u) & ~7u;
    return d;
}

// stages shared between entry points
// extrapolate: ghost faces no box of the level covers take the value of the
// box's nearest valid face (single-level calls on a finer level); off after
// the coarse-fine fill of FillPatchUedge, which has set them
int fill_face_ghosts(mx_adv_ctx* c, int lev, const mx_face_multifab* uedge,
                     bool extrapolate = true);
// fill: FillBoundary + FillUmacGhost afterwards (finest_level == 0); a
// hierarchy pass synchronises the levels instead
int addw0(mx_adv_ctx* c, int lev, const mx_face_multifab* uedge, const mx_face_multifab* w0mac,
          const mx_basestate* w0, double mult, bool fill = true);
// Maestro::FillPatchUedge over levels 0..nlev-1 (uedge[lev] per level)
int fill_patch_uedge(mx_adv_ctx* c, int nlev, const mx_face_multifab* uedge);
// the workspace VelPred's temporaries need on one level
size_t velpred_ws_bytes(const mx_adv_ctx* c, int lev);
int check_phys(const mx_adv_ctx* c);
int make_utrans(mx_adv_ctx* c, int lev, const mx_multifab* utilde, const mx_multifab* ufull,
                const mx_face_multifab* utrans, const mx_face_multifab* w0mac,
                const mx_multifab* w0_cart, bool fill = true);
// fused 3-D VelPred (ppm_type 1/2, forces not traced): the boxes of every
// listed level in one launch
struct VpLevel {
    int lev;
    const mx_multifab* utilde;
    const mx_multifab* ufull;
    const mx_face_multifab* utrans;
    const mx_face_multifab* umac;
    const mx_face_multifab* w0mac;
    const mx_multifab* force;
    const mx_multifab* w0_cart;
};
bool velpred_fused_ok(const mx_adv_ctx* c);
int velpred_fused(mx_adv_ctx* c, const std::vector<VpLevel>& levs);
int velpred(mx_adv_ctx* c, int lev, const mx_multifab* utilde, const mx_multifab* ufull,
            const mx_face_multifab* utrans, const mx_face_multifab* umac,
            const mx_face_multifab* w0mac, const mx_multifab* force, const mx_multifab* w0_cart);
// VelPred in two parts around MakeUtrans (the pass): setup uploads the box
// table, reserves the temporaries and runs the PPM kernel; MakeUtrans then
// reads the normal-direction traced states; finish runs the transverse and
// velocity kernels once utrans (with w0) is final
struct VpPlan {
    const void* dp = nullptr;  // device box table (VpBox)
    int nb = 0;
    long long maxpts = 0;
    int lev = 0;
    const mx_face_multifab* utrans = nullptr;
};
int velpred_setup(mx_adv_ctx* c, int lev, const mx_multifab* utilde, const mx_multifab* ufull,
                  const mx_face_multifab* utrans, const mx_face_multifab* umac,
                  const mx_face_multifab* w0mac, const mx_multifab* force,
                  const mx_multifab* w0_cart, VpPlan& plan, bool reserve = true);
int utrans_from_plan(mx_adv_ctx* c, const VpPlan& plan, bool fill = true);
int velpred_finish(mx_adv_ctx* c, const VpPlan& plan);
int make_edge_scal(mx_adv_ctx* c, int lev, const mx_multifab* state,
                   const mx_face_multifab* sedge, const mx_face_multifab* umac,
                   const mx_multifab* force, int is_vel, const mx_bcrec* bcs, int nbccomp,
                   int start_scomp, int start_bccomp, int num_comp, int is_conservative);
int make_rhox_flux(mx_adv_ctx* c, int lev, const mx_face_multifab* sflux,
                   const mx_multifab* etarhoflux, const mx_face_multifab* sedge,
                   const mx_face_multifab* umac, const mx_basestate* rho0_old,
                   const mx_basestate* rho0_edge_old, const mx_face_multifab* r0mac_old,
                   const mx_basestate* rho0_new, const mx_basestate* rho0_edge_new,
                   const mx_face_multifab* r0mac_new, const mx_basestate* rho0_predicted_edge,
                   const mx_basestate* w0, int start_comp, int num_comp, int zero_eta);
int make_rhoh_flux(mx_adv_ctx* c, int lev, const mx_face_multifab* sflux,
                   const mx_face_multifab* sedge, const mx_face_multifab* umac,
                   const mx_basestate* rho0_old, const mx_basestate* rho0_edge_old,
                   const mx_face_multifab* r0mac_old, const mx_basestate* rho0_new,
                   const mx_basestate* rho0_edge_new, const mx_face_multifab* r0mac_new,
                   const mx_basestate* rhoh0_old, const mx_basestate* rhoh0_edge_old,
                   const mx_face_multifab* rh0mac_old, const mx_basestate* rhoh0_new,
                   const mx_basestate* rhoh0_edge_new, const mx_face_multifab* rh0mac_new,
                   const mx_face_multifab* h0mac_old, const mx_face_multifab* h0mac_new);

// fused 3-D MakeEdgeScal (ppm_type 1/2, forces not traced); the _levels form
// runs the boxes of several levels in one launch
struct EdgeLevel {
    int lev;
    const mx_multifab* state;
    const mx_face_multifab* sedge;
    const mx_face_multifab* umac;
    const mx_multifab* force;
    int box = -1;  // one box of the level, or -1 for all
    // optional flux epilogue for state component flux_comp (MakeRhoHFlux's
    // planar branch of kind flux_kind: 1 umac (rhoh0_edge + sedge), 2 umac
    // sedge) into sflux, with the level's base states (cell, edge; old, new)
    int flux_comp = -1;
    int flux_kind = 0;
    const mx_face_multifab* sflux = nullptr;
    const double *bc_a = nullptr, *bc_b = nullptr, *be_a = nullptr, *be_b = nullptr;
};
bool fused_edge_scal_ok(const mx_adv_ctx* c, int is_conservative);
int make_edge_scal_fused_levels(mx_adv_ctx* c, const std::vector<EdgeLevel>& levs, int is_vel,
                                const mx_bcrec* bcs, int start_scomp, int start_bccomp,
                                int num_comp, int is_conservative);
// the same over a list of components (state component, BCRec index), all in
// one launch
int make_edge_scal_fused_list(mx_adv_ctx* c, const std::vector<EdgeLevel>& levs, int is_vel,
                              const mx_bcrec* bcs, const std::vector<int>& scomps,
                              const std::vector<int>& bccomps, int is_conservative);
int make_edge_scal_fused(mx_adv_ctx* c, int lev, const mx_multifab* state,
                         const mx_face_multifab* sedge, const mx_face_multifab* umac,
                         const mx_multifab* force, int is_vel, const mx_bcrec* bcs,
                         int start_scomp, int start_bccomp, int num_comp, int is_conservative);

// the pass (mx_adv_advance_pass and its multi-level / graph forms)
int pass_check(mx_adv_ctx* c, int lev, const mx_adv_pass_args* a);
int advance_pass_levels(mx_adv_ctx* c, int lev0, int nlev, const mx_adv_pass_args* a);

// flux registers (reflux_type 2)
int flux_reg_define(mx_adv_ctx* c, int lev, int ncomp);
void flux_reg_free(mx_adv_ctx* c, int lev);
int flux_reg_ops(mx_adv_ctx* c, int lev, const mx_face_multifab* sflux, bool rhox,
                 int start_comp, int num_comp);

void bc_setup_host(mx_adv_ctx* c);
int check_level(const mx_adv_ctx* c, int lev);
int check_mf(const mx_adv_ctx* c, int lev, const mx_multifab* m, int ncomp, int ngrow_needed,
             int nodal_dir);
int check_face(const mx_adv_ctx* c, int lev, const mx_face_multifab* f, int ncomp,
               int ngrow_needed);

}  // namespace mxh

#endif
