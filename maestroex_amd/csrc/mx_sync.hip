// mx_sync.hip -- the AMR restriction Maestro runs after each update:
// AverageDown (amrex::average_down) and AverageDownFaces
// (amrex::average_down_faces) between one pair of levels.
//
// The host intersects every coarse box with every coarsened fine box (cells,
// or the faces of one direction) and uploads the non-empty intersections as
// tasks; one launch covers them all, one thread per coarse point and
// component.  A coarse
// cell gets (1 / r^dim) times the sum of its r^dim fine cells, summed k
// outermost and i innermost; a coarse face the mean of the fine faces on
// it.  Coarsened fine cells never overlap; a coarse face shared by two
// coarsened fine boxes is written by both from their own copy of the fine
// face, which must hold one value (as it does after FillBoundary).
// HBM-bound: r^dim fine values in, one coarse value out.
#include "mx_internal.H"

using namespace mxd;
using namespace mxh;

namespace {

constexpr int NTS = 256;

struct SyncTask {
    int lo[3], hi[3];  // coarse points
    int t[3];          // fine points summed per coarse point in each direction
    int blk0;          // the task's first block
    double scale;      // 1 / (t0 t1 t2)
    DA crse, fine;
};

struct SyncArgs {
    int r[3];   // refinement ratio
    int comp, ncomp;
    int cells;  // 1: cell average (volfrac * sum), 0: face average (sum * facinv)
};

// grid: x over the blocks of all tasks (blk[x] names the task; no block is
// empty), z over components (one component per thread keeps every thread's
// loads independent of the store before them)
__global__ void __launch_bounds__(NTS) k_average_down(const SyncTask* __restrict__ tasks,
                                                      const int* __restrict__ blk, SyncArgs A) {
    const SyncTask T = tasks[blk[blockIdx.x]];
    int i, j, k;
    {
        const long long idx = (long long)(blockIdx.x - T.blk0) * NTS + threadIdx.x;
        if (!region_point(idx, T.lo, T.hi, i, j, k)) return;
    }
    const int ii = i * A.r[0], jj = j * A.r[1], kk = k * A.r[2];
    const int n = A.comp + (int)blockIdx.z;
    double s = 0.0;
    for (int kref = 0; kref < T.t[2]; ++kref)
        for (int jref = 0; jref < T.t[1]; ++jref)
            for (int iref = 0; iref < T.t[0]; ++iref)
                s += T.fine(ii + iref, jj + jref, kk + kref, n);
    T.crse(i, j, k, n) = A.cells ? T.scale * s : s * T.scale;
}

int floordiv(int a, int r) { return a >= 0 ? a / r : -((-a + r - 1) / r); }

HBox coarsen(const HBox& b, const int r[3]) {
    HBox o = b;
    for (int d = 0; d < 3; ++d) {
        o.lo[d] = floordiv(b.lo[d], r[d]);
        o.hi[d] = floordiv(b.hi[d], r[d]);
    }
    return o;
}

bool intersect(const HBox& a, const HBox& b, HBox& o) {
    o = a;
    for (int d = 0; d < 3; ++d) {
        o.lo[d] = std::max(a.lo[d], b.lo[d]);
        o.hi[d] = std::min(a.hi[d], b.hi[d]);
        if (o.lo[d] > o.hi[d]) return false;
    }
    return true;
}

// the refinement ratio between levels lev and lev + 1 (an integer in every
// direction, and the fine domain starting at lo * r), or 0
int ratio(const mx_adv_ctx* c, int lev, int r[3]) {
    for (int d = 0; d < 3; ++d) r[d] = 1;
    for (int d = 0; d < c->p.dim; ++d) {
        const int nc = c->p.domain_hi[lev][d] - c->p.domain_lo[lev][d] + 1;
        const int nf = c->p.domain_hi[lev + 1][d] - c->p.domain_lo[lev + 1][d] + 1;
        if (nc <= 0 || nf % nc != 0 || nf / nc < 1) return 0;
        r[d] = nf / nc;
        if (c->p.domain_lo[lev + 1][d] != c->p.domain_lo[lev][d] * r[d]) return 0;
    }
    return 1;
}

// the tasks of one cell or face direction (d < 0: cells); every fine point
// a task reads is checked to lie in the fine FAB
int tasks_for(const mx_adv_ctx* c, int lev, int d, const mx_multifab* fine,
              const mx_multifab* crse, const int r[3], int ncomp_needed,
              std::vector<SyncTask>& out) {
    const auto& fg = c->grids[lev + 1];
    const auto& cg = c->grids[lev];
    for (size_t f = 0; f < fg.size(); ++f) {
        HBox cf = coarsen(hbox(fg[f]), r);
        if (d >= 0) cf = nodal(cf, d);
        for (size_t b = 0; b < cg.size(); ++b) {
            HBox cb = hbox(cg[b]);
            if (d >= 0) cb = nodal(cb, d);
            HBox reg;
            if (!intersect(cf, cb, reg)) continue;
            // the fine points read: [lo * r, hi * r + r - 1] (cells) or the
            // faces at hi * r in direction d
            HBox fr = reg;
            for (int q = 0; q < 3; ++q) {
                fr.lo[q] = reg.lo[q] * r[q];
                fr.hi[q] = reg.hi[q] * r[q] + ((q == d) ? 0 : r[q] - 1);
            }
            if (!covers(fine->fab[f], fr, ncomp_needed) || !covers(crse->fab[b], reg, ncomp_needed))
                return MX_ERR_BOX;
            SyncTask T;
            for (int q = 0; q < 3; ++q) {
                T.lo[q] = reg.lo[q];
                T.hi[q] = reg.hi[q];
                T.t[q] = q == d ? 1 : r[q];
            }
            T.scale = 1.0 / double(T.t[0] * T.t[1] * T.t[2]);
            T.crse = to_da(crse->fab[b]);
            T.fine = to_da(fine->fab[f]);
            T.blk0 = 0;
            out.push_back(T);
        }
    }
    return MX_OK;
}

int launch_tasks(mx_adv_ctx* c, std::vector<SyncTask>& tasks, const SyncArgs& A) {
    if (tasks.empty()) return MX_OK;
    if (A.ncomp > 65535) return MX_ERR_UNSUPPORTED;
    std::vector<int> blk;
    for (size_t q = 0; q < tasks.size(); ++q) {
        const SyncTask& T = tasks[q];
        long long np = 1;
        for (int d = 0; d < 3; ++d) np *= (long long)(T.hi[d] - T.lo[d] + 1);
        const long long nb = (np + NTS - 1) / NTS;
        if ((long long)blk.size() + nb > 0x7fffffffLL) return MX_ERR_UNSUPPORTED;
        tasks[q].blk0 = (int)blk.size();
        blk.insert(blk.end(), (size_t)nb, (int)q);
    }
    Table tab(c);
    const size_t ot = tab.add(tasks);
    const size_t ob = tab.add(blk);
    int rc = tab.upload();
    if (rc) return rc;
    MX_LAUNCH(c, "k_average_down", k_average_down,
              dim3((unsigned)blk.size(), 1, (unsigned)A.ncomp), dim3(NTS), tab.dev<SyncTask>(ot),
              tab.dev<int>(ob), A);
    return MX_OK;
}

}  // namespace

extern "C" {

int mx_adv_average_down(mx_adv_ctx* c, int lev, const mx_multifab* fine, mx_multifab* crse,
                        int comp, int ncomp) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    if (lev + 1 > c->p.finest_level || c->grids[lev + 1].empty()) return MX_ERR_ARG;
    if (!fine || !crse || comp < 0 || ncomp < 1) return MX_ERR_ARG;
    if ((int)c->grids[lev + 1].size() != fine->nfab || (int)c->grids[lev].size() != crse->nfab)
        return MX_ERR_ARG;
    int r[3];
    if (!ratio(c, lev, r)) return MX_ERR_ARG;
    std::vector<SyncTask> tasks;
    if ((rc = tasks_for(c, lev, -1, fine, crse, r, comp + ncomp, tasks))) return rc;
    SyncArgs A;
    for (int d = 0; d < 3; ++d) A.r[d] = r[d];
    A.comp = comp;
    A.ncomp = ncomp;
    A.cells = 1;
    return launch_tasks(c, tasks, A);
}

int mx_adv_average_down_faces(mx_adv_ctx* c, int lev, const mx_face_multifab* fine,
                              mx_face_multifab* crse) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    if (lev + 1 > c->p.finest_level || c->grids[lev + 1].empty()) return MX_ERR_ARG;
    if (!fine || !crse) return MX_ERR_ARG;
    int r[3];
    if (!ratio(c, lev, r)) return MX_ERR_ARG;
    // every direction's faces in one launch (one component count for all)
    std::vector<SyncTask> tasks;
    int nc = -1;
    for (int d = 0; d < c->p.dim; ++d) {
        const mx_multifab* f = &fine->dir[d];
        mx_multifab* cm = &crse->dir[d];
        if ((int)c->grids[lev + 1].size() != f->nfab || (int)c->grids[lev].size() != cm->nfab ||
            !f->fab || !cm->fab)
            return MX_ERR_ARG;
        if (nc < 0) nc = cm->fab[0].ncomp;
        if (cm->fab[0].ncomp != nc) return MX_ERR_ARG;
        if ((rc = tasks_for(c, lev, d, f, cm, r, nc, tasks))) return rc;
    }
    SyncArgs A;
    for (int q = 0; q < 3; ++q) A.r[q] = r[q];
    A.comp = 0;
    A.ncomp = nc;
    A.cells = 0;
    return launch_tasks(c, tasks, A);
}

}  // extern "C"
