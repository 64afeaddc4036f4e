// mx_amr.hip -- coarse-fine synchronisation of face data between AMR levels:
// Maestro::FillPatchUedge (MaestroFillData.cpp:455-833).
//
// On a level lev > 0 every ghost face of a fine box that no box of the level
// holds gets its value from the coarser level and the level's own valid
// faces.  The reference builds buffers over the level's boundary cells
// (coarsened and refined again), fills them piecewise constant from the
// coarse faces, overwrites them with the fine valid faces, interpolates
// linearly in the normal direction between coarse-aligned faces and copies
// them into the ghost faces.  Each buffer value depends only on the face's
// position, so one thread per ghost face computes it directly:
//
//   r = refinement ratio, a = f[dir] mod r (f the fine face index);
//   src(g) = the fine valid face g of a box of the level (later boxes win),
//            else the coarse face (g[dir] / r, floor(g_t / r)) of a coarse
//            FAB, ghosts included (later boxes win), else 1e200;
//   a == 0: value = src(f);
//   a != 0: g0 = (f[dir] - a, transverse indices rounded down to multiples of
//           r), g1 = g0 + r in dir; value = src(g0) + (src(g1) - src(g0)) * a / r.
//
// The ghost faces other boxes of the level hold are then overwritten by
// FillBoundary (with periodicity) and those behind physical boundaries by
// FillUmacGhost, as in the reference (:824-831).
#include <algorithm>

#include "mx_internal.H"

using namespace mxd;
using namespace mxh;

namespace {

struct CFSlab {
    DA dst;
    int lo[3], hi[3];
    int f_off, f_n;  // fine candidates
    int c_off, c_n;  // coarse candidates
};

struct CFFine {
    DA src;
    int vlo[3], vhi[3];  // valid faces of the fine box
};

struct CFCoarse {
    DA src;
    int glo[3], ghi[3];  // faces of the coarse FAB (valid grown by one)
};

struct CFArgs {
    int dir;
    int r[3];
};

__device__ __forceinline__ int fdiv(int a, int r) { return a >= 0 ? a / r : -((-a + r - 1) / r); }

__device__ __forceinline__ double cf_src(const CFSlab& t, const CFFine* __restrict__ fc,
                                         const CFCoarse* __restrict__ cc, const CFArgs& A,
                                         const int g[3]) {
    for (int q = t.f_n - 1; q >= 0; --q) {
        const CFFine& f = fc[t.f_off + q];
        if (g[0] >= f.vlo[0] && g[0] <= f.vhi[0] && g[1] >= f.vlo[1] && g[1] <= f.vhi[1] &&
            g[2] >= f.vlo[2] && g[2] <= f.vhi[2])
            return f.src(g[0], g[1], g[2]);
    }
    const int c0 = fdiv(g[0], A.r[0]), c1 = fdiv(g[1], A.r[1]), c2 = fdiv(g[2], A.r[2]);
    for (int q = t.c_n - 1; q >= 0; --q) {
        const CFCoarse& c = cc[t.c_off + q];
        if (c0 >= c.glo[0] && c0 <= c.ghi[0] && c1 >= c.glo[1] && c1 <= c.ghi[1] &&
            c2 >= c.glo[2] && c2 <= c.ghi[2])
            return c.src(c0, c1, c2);
    }
    return 1.e200;
}

__global__ void __launch_bounds__(256) k_cf_fill_faces(const CFSlab* __restrict__ slabs,
                                                       const CFFine* __restrict__ fc,
                                                       const CFCoarse* __restrict__ cc,
                                                       CFArgs A) {
    const CFSlab& t = slabs[blockIdx.y];
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!region_point(idx, t.lo, t.hi, i, j, k)) return;
    const int f[3] = {i, j, k};
    const int d = A.dir;
    const int rd = A.r[d];
    const int a = f[d] - fdiv(f[d], rd) * rd;
    double v;
    if (a == 0) {
        v = cf_src(t, fc, cc, A, f);
    } else {
        int g0[3], g1[3];
        for (int q = 0; q < 3; ++q) g0[q] = q == d ? f[q] - a : fdiv(f[q], A.r[q]) * A.r[q];
        for (int q = 0; q < 3; ++q) g1[q] = g0[q] + (q == d ? rd : 0);
        const double s0 = cf_src(t, fc, cc, A, g0);
        const double s1 = cf_src(t, fc, cc, A, g1);
        const double df = s1 - s0;
        v = s0 + df * double(a) / double(rd);
    }
    t.dst(i, j, k) = v;
}

int hfdiv(int a, int r) { return a >= 0 ? a / r : -((-a + r - 1) / r); }

bool boxes_meet(const HBox& a, const HBox& b) {
    for (int q = 0; q < 3; ++q)
        if (std::max(a.lo[q], b.lo[q]) > std::min(a.hi[q], b.hi[q])) return false;
    return true;
}

// the ghost-face slabs of every box of fine level lev (normal dir) with their
// candidate fine and coarse source boxes
struct CFPlan {
    struct Slab {
        int box;
        int lo[3], hi[3];
        int f_off, f_n, c_off, c_n;
    };
    std::vector<Slab> slabs;
    std::vector<int> fine;    // fine box indices
    std::vector<int> coarse;  // coarse box indices
};

CFPlan cf_plan(const mx_adv_ctx* c, int lev, int dir, const int r[3]) {
    const int dim = c->p.dim;
    const auto& fg = c->grids[lev];
    const auto& cg = c->grids[lev - 1];
    CFPlan plan;
    for (size_t b = 0; b < fg.size(); ++b) {
        const HBox valid = nodal(hbox(fg[b]), dir);
        const HBox fb = grow(valid, dim, 1);
        for (int d = 0; d < dim; ++d)
            for (int side = 0; side < 2; ++side) {
                CFPlan::Slab sl;
                sl.box = (int)b;
                for (int q = 0; q < 3; ++q) {
                    sl.lo[q] = q < d ? valid.lo[q] : fb.lo[q];
                    sl.hi[q] = q < d ? valid.hi[q] : fb.hi[q];
                }
                if (side == 0)
                    sl.hi[d] = valid.lo[d] - 1;
                else
                    sl.lo[d] = valid.hi[d] + 1;
                if (sl.lo[d] > sl.hi[d]) continue;
                // every face the slab's values read: within r of the slab
                HBox reach;
                for (int q = 0; q < 3; ++q) {
                    reach.lo[q] = sl.lo[q] - (q < dim ? r[q] : 0);
                    reach.hi[q] = sl.hi[q] + (q < dim ? r[q] : 0);
                }
                sl.f_off = (int)plan.fine.size();
                for (size_t s = 0; s < fg.size(); ++s)
                    if (boxes_meet(reach, nodal(hbox(fg[s]), dir))) plan.fine.push_back((int)s);
                sl.f_n = (int)plan.fine.size() - sl.f_off;
                HBox creach;
                for (int q = 0; q < 3; ++q) {
                    creach.lo[q] = hfdiv(reach.lo[q], r[q]);
                    creach.hi[q] = hfdiv(reach.hi[q], r[q]);
                }
                sl.c_off = (int)plan.coarse.size();
                for (size_t B = 0; B < cg.size(); ++B)
                    if (boxes_meet(creach, grow(nodal(hbox(cg[B]), dir), dim, 1)))
                        plan.coarse.push_back((int)B);
                sl.c_n = (int)plan.coarse.size() - sl.c_off;
                plan.slabs.push_back(sl);
            }
    }
    return plan;
}

int level_ratio(const mx_adv_ctx* c, int lev, int r[3]) {
    for (int q = 0; q < 3; ++q) {
        r[q] = 1;
        if (q >= c->p.dim) continue;
        const int nc = c->p.domain_hi[lev - 1][q] - c->p.domain_lo[lev - 1][q] + 1;
        const int nf = c->p.domain_hi[lev][q] - c->p.domain_lo[lev][q] + 1;
        if (nc <= 0 || nf % nc != 0 || nf / nc < 1) return MX_ERR_ARG;
        r[q] = nf / nc;
    }
    return MX_OK;
}

// the coarse-fine fill of one fine level, every direction
int cf_fill(mx_adv_ctx* c, int lev, const mx_face_multifab* crse, const mx_face_multifab* fine) {
    int r[3];
    int rc = level_ratio(c, lev, r);
    if (rc) return rc;
    const int dim = c->p.dim;
    const auto& fg = c->grids[lev];
    const auto& cg = c->grids[lev - 1];
    for (int d = 0; d < dim; ++d) {
        const CFPlan plan = cf_plan(c, lev, d, r);
        if (plan.slabs.empty()) continue;
        std::vector<CFSlab> slabs;
        std::vector<CFFine> fc;
        std::vector<CFCoarse> cc;
        for (int s : plan.fine) {
            CFFine f;
            f.src = to_da(fine->dir[d].fab[s]);
            const HBox v = nodal(hbox(fg[s]), d);
            for (int q = 0; q < 3; ++q) {
                f.vlo[q] = v.lo[q];
                f.vhi[q] = v.hi[q];
            }
            fc.push_back(f);
        }
        for (int B : plan.coarse) {
            CFCoarse g;
            g.src = to_da(crse->dir[d].fab[B]);
            const HBox v = grow(nodal(hbox(cg[B]), d), dim, 1);
            for (int q = 0; q < 3; ++q) {
                g.glo[q] = v.lo[q];
                g.ghi[q] = v.hi[q];
            }
            cc.push_back(g);
        }
        long long maxpts = 0;
        for (const auto& sl : plan.slabs) {
            CFSlab t;
            t.dst = to_da(fine->dir[d].fab[sl.box]);
            HBox reg;
            for (int q = 0; q < 3; ++q) {
                t.lo[q] = reg.lo[q] = sl.lo[q];
                t.hi[q] = reg.hi[q] = sl.hi[q];
            }
            t.f_off = sl.f_off;
            t.f_n = sl.f_n;
            t.c_off = sl.c_off;
            t.c_n = sl.c_n;
            maxpts = std::max(maxpts, npts(reg));
            slabs.push_back(t);
        }
        CFArgs A;
        A.dir = d;
        for (int q = 0; q < 3; ++q) A.r[q] = r[q];
        Table tab(c);
        const size_t os = tab.add(slabs);
        const size_t of = tab.add(fc);
        const size_t oc = tab.add(cc);
        if ((rc = tab.upload())) return rc;
        const int block = 256;
        MX_LAUNCH(c, "k_cf_fill_faces", k_cf_fill_faces,
                  grid_for(maxpts, (int)slabs.size(), 1, block), dim3(block),
                  tab.dev<CFSlab>(os), tab.dev<CFFine>(of), tab.dev<CFCoarse>(oc), A);
    }
    return MX_OK;
}

}  // namespace

namespace mxh {

int fill_patch_uedge(mx_adv_ctx* c, int nlev, const mx_face_multifab* uedge) {
    for (int lev = 0; lev < nlev; ++lev) {
        int rc;
        if (lev > 0 && (rc = cf_fill(c, lev, &uedge[lev - 1], &uedge[lev]))) return rc;
        if ((rc = fill_face_ghosts(c, lev, &uedge[lev], nlev == 1))) return rc;
    }
    return MX_OK;
}

}  // namespace mxh

extern "C" {

int mx_adv_fill_patch_uedge(mx_adv_ctx* c, int nlev, mx_face_multifab* uedge) {
    if (!c || !uedge || nlev < 1 || nlev - 1 > c->p.finest_level) return MX_ERR_ARG;
    for (int lev = 0; lev < nlev; ++lev) {
        int rc = check_level(c, lev);
        if (rc) return rc;
        if ((rc = check_face(c, lev, &uedge[lev], 1, 1))) return rc;
    }
    return fill_patch_uedge(c, nlev, uedge);
}

}  // extern "C"
