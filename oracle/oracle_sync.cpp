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

}  // namespace mxo
