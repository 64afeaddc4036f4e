// oracle_sync.cpp -- CPU restatement of the AMR restriction Maestro calls
// after each update: AverageDown (MaestroFillData.cpp:145-154, a loop over
// amrex::average_down) and AverageDownFaces (:156-172, a loop over
// amrex::average_down_faces), for one pair of levels.
//
// TEST INFRASTRUCTURE ONLY (see oracle_core.H).  amrex::average_down and
// average_down_faces live in AMReX (AMReX_MultiFabUtil), an un-vendored
// submodule of the reference (external/amrex is empty and its pinned commit
// is not recorded).  This restates their published algorithm: a coarse cell
// covered by the coarsened fine BoxArray gets (1 / r^dim) times the sum of
// its r^dim fine cells, summed k outermost and i innermost; a coarse face
// covered by the coarsened fine faces gets 1 / (r^(dim-1)) times the sum of
// the fine faces on it, in the same order.  Coarse points not covered keep
// their values; coarse ghosts are not touched.
#include "oracle_core.H"

namespace mxo {

namespace {

int ratio_of(const Ctx& c, int lev, int d) {
    const IBox a = c.domain(lev), b = c.domain(lev + 1);
    return (b.hi[d] - b.lo[d] + 1) / (a.hi[d] - a.lo[d] + 1);
}

int floordiv(int a, int r) { return a >= 0 ? a / r : -((-a + r - 1) / r); }

IBox coarsen(const IBox& b, const int r[3]) {
    IBox o;
    for (int d = 0; d < 3; ++d) {
        o.lo[d] = floordiv(b.lo[d], r[d]);
        o.hi[d] = floordiv(b.hi[d], r[d]);
    }
    return o;
}

bool intersect(const IBox& a, const IBox& b, IBox& o) {
    for (int d = 0; d < 3; ++d) {
        o.lo[d] = a.lo[d] > b.lo[d] ? a.lo[d] : b.lo[d];
        o.hi[d] = a.hi[d] < b.hi[d] ? a.hi[d] : b.hi[d];
        if (o.lo[d] > o.hi[d]) return false;
    }
    return true;
}

}  // namespace

// amrex::average_down(fine, crse, ..., scomp, ncomp, ratio) between levels
// lev + 1 and lev
int AverageDown(const Ctx& c, int lev, const std::vector<Arr>& fine,
                const std::vector<Arr>& crse, int comp, int ncomp) {
    const int dim = c.dim();
    int r[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) r[d] = ratio_of(c, lev, d);
    const double volfrac = 1.0 / double(r[0] * r[1] * r[2]);
    const std::vector<IBox>& fg = c.grids[lev + 1];
    const std::vector<IBox>& cg = c.grids[lev];
    for (size_t f = 0; f < fg.size(); ++f) {
        const IBox cf = coarsen(fg[f], r);
        const Arr& fa = fine[f];
        for (size_t b = 0; b < cg.size(); ++b) {
            IBox reg;
            if (!intersect(cf, cg[b], reg)) continue;
            const Arr& ca = crse[b];
            par_for(reg, [&](int i, int j, int k) {
                const int ii = i * r[0], jj = j * r[1], kk = k * r[2];
                for (int n = comp; n < comp + ncomp; ++n) {
                    double s = 0.0;
                    for (int kref = 0; kref < r[2]; ++kref)
                        for (int jref = 0; jref < r[1]; ++jref)
                            for (int iref = 0; iref < r[0]; ++iref)
                                s += fa(ii + iref, jj + jref, kk + kref, n);
                    ca(i, j, k, n) = volfrac * s;
                }
            });
        }
    }
    return MX_OK;
}

// amrex::average_down_faces(fine, crse, ratio) between levels lev + 1 and lev
int AverageDownFaces(const Ctx& c, int lev, const std::vector<Arr> fine[3],
                     const std::vector<Arr> crse[3]) {
    const int dim = c.dim();
    int r[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) r[d] = ratio_of(c, lev, d);
    const std::vector<IBox>& fg = c.grids[lev + 1];
    const std::vector<IBox>& cg = c.grids[lev];
    for (int d = 0; d < dim; ++d) {
        int t[3] = {r[0], r[1], r[2]};
        t[d] = 1;  // the fine faces on a coarse face: the transverse refinements
        const double facinv = 1.0 / double(t[0] * t[1] * t[2]);
        for (size_t f = 0; f < fg.size(); ++f) {
            const IBox cf = nodal(coarsen(fg[f], r), d);
            const Arr& fa = fine[d][f];
            for (size_t b = 0; b < cg.size(); ++b) {
                IBox reg;
                if (!intersect(cf, nodal(cg[b], d), reg)) continue;
                const Arr& ca = crse[d][b];
                par_for(reg, [&](int i, int j, int k) {
                    const int ii = i * r[0], jj = j * r[1], kk = k * r[2];
                    for (int n = 0; n < ca.nc; ++n) {
                        double s = 0.0;
                        for (int kref = 0; kref < t[2]; ++kref)
                            for (int jref = 0; jref < t[1]; ++jref)
                                for (int iref = 0; iref < t[0]; ++iref)
                                    s += fa(ii + iref, jj + jref, kk + kref, n);
                        ca(i, j, k, n) = s * facinv;
                    }
                });
            }
        }
    }
    return MX_OK;
}

// ---------------------------------------------------------------------------
// Maestro::FillPatchUedge on a level lev > 0, the part that fills the fine
// ghost faces from level lev - 1 (MaestroFillData.cpp:466-821).  The periodic
// FillBoundary and FillUmacGhost the reference runs next on every level are
// FillBoundaryFace (without its uncovered-face extrapolation) and
// FillUmacGhost.
//
// Steps, per direction dir:
//   * the boundary cells of the level (amrex::GetBndryCells with one ghost:
//     cells of the grown boxes no box of the level holds), coarsened and
//     refined again (so that they cover whole coarse cells);
//   * per such box, a coarse buffer on its coarsened faces set to 1e200 and
//     copied from the coarse faces, ghosts included (ParallelCopy with one
//     source ghost; later source boxes overwrite earlier ones), and a fine
//     buffer set to 1e200 and filled piecewise constant on the fine faces
//     that lie on coarse faces (:539-628);
//   * the fine buffer takes the level's valid fine faces where they exist
//     (ParallelCopy without ghosts, :633);
//   * the faces between two coarse-aligned faces are interpolated linearly in
//     dir from the aligned faces of the coarse-aligned transverse row,
//     piecewise constant across (:637-808);
//   * each fine FAB's ghost faces take the buffers' values (later buffers
//     overwrite earlier ones); its valid faces are kept (:810-820).
namespace {

IBox refine(const IBox& b, const int r[3]) {
    IBox o;
    for (int d = 0; d < 3; ++d) {
        o.lo[d] = b.lo[d] * r[d];
        o.hi[d] = (b.hi[d] + 1) * r[d] - 1;
    }
    return o;
}

bool empty_box(const IBox& b) {
    for (int d = 0; d < 3; ++d)
        if (b.lo[d] > b.hi[d]) return true;
    return false;
}

// a \ b as disjoint boxes
void box_diff(const IBox& a, const IBox& b, std::vector<IBox>& out) {
    IBox in;
    if (!intersect(a, b, in)) {
        out.push_back(a);
        return;
    }
    IBox rest = a;
    for (int d = 0; d < 3; ++d) {
        if (rest.lo[d] < in.lo[d]) {
            IBox s = rest;
            s.hi[d] = in.lo[d] - 1;
            out.push_back(s);
            rest.lo[d] = in.lo[d];
        }
        if (rest.hi[d] > in.hi[d]) {
            IBox s = rest;
            s.lo[d] = in.hi[d] + 1;
            out.push_back(s);
            rest.hi[d] = in.hi[d];
        }
    }
}

// amrex::GetBndryCells(ba, ngrow): cells of the grown boxes outside every box
std::vector<IBox> bndry_cells(const std::vector<IBox>& ba, int dim, int ngrow) {
    std::vector<IBox> out;
    for (const IBox& b : ba) {
        std::vector<IBox> pieces;
        box_diff(grow(b, dim, ngrow), b, pieces);
        for (const IBox& o : ba) {
            std::vector<IBox> next;
            for (const IBox& q : pieces) box_diff(q, o, next);
            pieces.swap(next);
        }
        for (const IBox& q : pieces)
            if (!empty_box(q)) out.push_back(q);
    }
    return out;
}

struct Buf {
    IBox bx;
    std::vector<double> v;
    long long nx = 0, ny = 0;
    void define(const IBox& b, double val) {
        bx = b;
        nx = b.hi[0] - b.lo[0] + 1;
        ny = b.hi[1] - b.lo[1] + 1;
        v.assign((size_t)(nx * ny * (b.hi[2] - b.lo[2] + 1)), val);
    }
    double& operator()(int i, int j, int k) {
        return v[(size_t)((i - bx.lo[0]) + nx * ((j - bx.lo[1]) + ny * (k - bx.lo[2])))];
    }
};

}  // namespace

int FillPatchUedgeCF(const Ctx& c, int lev, const std::vector<Arr> crse[3],
                     const std::vector<Arr> fine[3]) {
    if (lev < 1) return MX_ERR_ARG;
    const int dim = c.dim();
    int r[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) r[d] = ratio_of(c, lev - 1, d);
    const std::vector<IBox>& fg = c.grids[lev];
    const std::vector<IBox>& cg = c.grids[lev - 1];
    // boundary cells, coarsened, then refined so fine and coarse cover the
    // same region
    std::vector<IBox> cb;
    for (const IBox& b : bndry_cells(fg, dim, 1)) cb.push_back(coarsen(b, r));
    for (int dir = 0; dir < dim; ++dir) {
        std::vector<Buf> bufs(cb.size());
        for (size_t q = 0; q < cb.size(); ++q) {
            const IBox cbox = nodal(cb[q], dir);
            const IBox fbox = nodal(refine(cb[q], r), dir);
            Buf crs;
            crs.define(cbox, 1.e200);
            for (size_t B = 0; B < cg.size(); ++B) {
                IBox in;
                if (!intersect(cbox, grow(nodal(cg[B], dir), dim, 1), in)) continue;
                const Arr& a = crse[dir][B];
                for (int k = in.lo[2]; k <= in.hi[2]; ++k)
                    for (int j = in.lo[1]; j <= in.hi[1]; ++j)
                        for (int i = in.lo[0]; i <= in.hi[0]; ++i) crs(i, j, k) = a(i, j, k);
            }
            Buf& fs = bufs[q];
            fs.define(fbox, 1.e200);
            // piecewise constant on the fine faces of each coarse face
            const int ra = dir == 0 ? 1 : 0, rb = dir == 2 ? 1 : 2;  // transverse directions
            for (int k = cbox.lo[2]; k <= cbox.hi[2]; ++k)
                for (int j = cbox.lo[1]; j <= cbox.hi[1]; ++j)
                    for (int i = cbox.lo[0]; i <= cbox.hi[0]; ++i) {
                        const int f0[3] = {i * r[0], j * r[1], k * r[2]};
                        const int na = ra < dim ? r[ra] : 1, nb = rb < dim ? r[rb] : 1;
                        for (int P = 0; P < nb; ++P)
                            for (int L = 0; L < na; ++L) {
                                int f[3] = {f0[0], f0[1], f0[2]};
                                f[ra] += L;
                                f[rb] += P;
                                fs(f[0], f[1], f[2]) = crs(i, j, k);
                            }
                    }
            // the level's valid fine faces where they exist
            for (size_t b = 0; b < fg.size(); ++b) {
                IBox in;
                if (!intersect(fbox, nodal(fg[b], dir), in)) continue;
                const Arr& a = fine[dir][b];
                for (int k = in.lo[2]; k <= in.hi[2]; ++k)
                    for (int j = in.lo[1]; j <= in.hi[1]; ++j)
                        for (int i = in.lo[0]; i <= in.hi[0]; ++i) fs(i, j, k) = a(i, j, k);
            }
            // linear in dir between aligned faces, piecewise constant across
            const IBox& f = fbox;
            const int rd = r[dir];
            int step[3] = {r[0], r[1], r[2]};
            for (int d = dim; d < 3; ++d) step[d] = 1;
            int top[3] = {f.hi[0], f.hi[1], f.hi[2]};
            top[dir] -= rd;
            for (int k = f.lo[2]; k <= top[2]; k += step[2])
                for (int j = f.lo[1]; j <= top[1]; j += step[1])
                    for (int i = f.lo[0]; i <= top[0]; i += step[0]) {
                        const int g0[3] = {i, j, k};
                        int g1[3] = {i, j, k};
                        g1[dir] += rd;
                        const double df = fs(g1[0], g1[1], g1[2]) - fs(g0[0], g0[1], g0[2]);
                        for (int M = 1; M < rd; ++M) {
                            const double val = fs(g0[0], g0[1], g0[2]) + df * double(M) / double(rd);
                            const int na = ra < dim ? step[ra] : 1, nb = rb < dim ? step[rb] : 1;
                            for (int P = 0; P < nb; ++P) {
                                const int pb = g0[rb] + P;
                                if (pb > f.hi[rb]) break;
                                for (int L = 0; L < na; ++L) {
                                    const int pa = g0[ra] + L;
                                    if (pa > f.hi[ra]) break;
                                    int t[3] = {g0[0], g0[1], g0[2]};
                                    t[dir] += M;
                                    t[ra] = pa;
                                    t[rb] = pb;
                                    fs(t[0], t[1], t[2]) = val;
                                }
                            }
                        }
                    }
        }
        // ghost faces of each fine FAB from the buffers; valid faces kept
        for (size_t b = 0; b < fg.size(); ++b) {
            const IBox valid = nodal(fg[b], dir);
            const IBox gb = grow(valid, dim, 1);
            const Arr& a = fine[dir][b];
            for (size_t q = 0; q < cb.size(); ++q) {
                IBox in;
                if (!intersect(gb, bufs[q].bx, in)) continue;
                for (int k = in.lo[2]; k <= in.hi[2]; ++k)
                    for (int j = in.lo[1]; j <= in.hi[1]; ++j)
                        for (int i = in.lo[0]; i <= in.hi[0]; ++i) {
                            if (i >= valid.lo[0] && i <= valid.hi[0] && j >= valid.lo[1] &&
                                j <= valid.hi[1] && k >= valid.lo[2] && k <= valid.hi[2])
                                continue;
                            a(i, j, k) = bufs[q](i, j, k);
                        }
            }
        }
    }
    return MX_OK;
}

}  // namespace mxo
