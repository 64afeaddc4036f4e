// oracle_fill.cpp -- CPU restatement of Maestro::FillPatch at level 0
// (MaestroFillData.cpp:24-47): amrex::FillPatchSingleLevel with Maestro's
// physical-boundary functions ScalarFill / VelFill (MaestroBCFill.cpp:8-28)
// and FillExtBC (:30-171).
//
// TEST INFRASTRUCTURE ONLY (see oracle_core.H).  FillPatchSingleLevel,
// PhysBCFunct and fab_filcc live in AMReX, an un-vendored submodule of the
// reference (external/amrex is empty and its pinned commit is not recorded).
// This restates their published algorithm:
//   1. ParallelCopy: every cell of the destination's grown box that lies in a
//      source valid box, or in a periodic image of one, takes the source
//      value (scomp.. -> dcomp..);
//   2. for every destination FAB whose grown box is not inside the domain
//      grown in the periodic directions, fab_filcc per component and BCRec:
//      below domlo in direction d,
//        foextrap      q(i) = q(ilo)
//        hoextrap      i <  ilo-1: q(ilo);  i == ilo-1: 1/8 (15 q(ilo) -
//                      10 q(ilo+1) + 3 q(ilo+2)) when the FAB reaches ilo+2,
//                      else 1/2 (3 q(ilo) - q(ilo+1))
//        reflect_even  q(i) =  q(2 ilo - i - 1)
//        reflect_odd   q(i) = -q(2 ilo - i - 1)
//        int_dir, ext_dir: untouched
//      and the mirror image above domhi;
//   3. FillExtBC: with ext_dir on a side, the cells beyond it are set to 0
//      for Pi and for velocity; for any other scalar the reference aborts
//      (here MX_ERR_BC, before anything is written).
// fab_filcc visits cells in one loop and applies the x, y and z rules to
// each in turn, so a cell outside the domain in two non-periodic directions
// (a corner) reads neighbours the same loop may not have filled yet.  Here
// every direction is one pass over all cells (x, then y, then z), so a
// corner takes the last direction's rule applied to values the earlier
// passes filled.  Cells outside the domain in one direction only, and every
// corner between a periodic and a non-periodic direction, do not depend on
// the order.
#include "oracle_core.H"

namespace mxo {

namespace {

bool inside(const IBox& a, const IBox& b) {  // a inside b
    for (int d = 0; d < 3; ++d)
        if (a.lo[d] < b.lo[d] || a.hi[d] > b.hi[d]) return false;
    return true;
}

bool intersect(const IBox& a, const IBox& b, IBox& o) {
    for (int d = 0; d < 3; ++d) {
        o.lo[d] = a.lo[d] > b.lo[d] ? a.lo[d] : b.lo[d];
        o.hi[d] = a.hi[d] < b.hi[d] ? a.hi[d] : b.hi[d];
        if (o.lo[d] > o.hi[d]) return false;
    }
    return true;
}

IBox fab_box(const Arr& a) {
    IBox b;
    for (int d = 0; d < 3; ++d) {
        b.lo[d] = a.b[d];
        b.hi[d] = a.e[d] - 1;
    }
    return b;
}

}  // namespace

int FillPatchSingleLevel(const Ctx& c, int lev, const std::vector<Arr>& mf,
                         const std::vector<Arr>& src, int scomp, int dcomp, int ncomp,
                         const mx_bcrec* bcs, bool is_vel) {
    const int dim = c.dim();
    const IBox dom = c.domain(lev);
    const std::vector<IBox>& g = c.grids[lev];
    // BCRec values and the FillExtBC abort, before anything is written
    for (int n = 0; n < ncomp; ++n)
        for (int d = 0; d < dim; ++d)
            for (int side = 0; side < 2; ++side) {
                const int bc = side ? bcs[n].hi[d] : bcs[n].lo[d];
                if (bc < MX_BC_REFLECT_ODD || bc > MX_BC_HOEXTRAP) return MX_ERR_BC;
                if (bc != MX_BC_EXT_DIR || is_vel || dcomp + n == c.p.comp_pi) continue;
                for (const Arr& a : mf) {
                    const IBox fb = fab_box(a);
                    if (side ? fb.hi[d] > dom.hi[d] : fb.lo[d] < dom.lo[d]) return MX_ERR_BC;
                }
            }
    // the domain grown by the ghost width in the periodic directions
    IBox gdom = dom;
    for (int d = 0; d < dim; ++d)
        if (c.p.is_periodic[d] && !mf.empty() && !g.empty()) {
            const int ng = g[0].lo[d] - fab_box(mf[0]).lo[d];
            gdom.lo[d] -= ng;
            gdom.hi[d] += ng;
        }
    // every cell a boundary rule reads lies in the FAB
    for (const Arr& a : mf) {
        const IBox fb = fab_box(a);
        if (inside(fb, gdom)) continue;
        for (int d = 0; d < dim; ++d) {
            const int reach_lo = 2 * dom.lo[d] - fb.lo[d] - 1;  // reflection of the lowest cell
            const int reach_hi = 2 * dom.hi[d] - fb.hi[d] + 1;
            if (fb.lo[d] < dom.lo[d] && (fb.hi[d] < dom.lo[d] + 1 || fb.hi[d] < reach_lo))
                return MX_ERR_BOX;
            if (fb.hi[d] > dom.hi[d] && (fb.lo[d] > dom.hi[d] - 1 || fb.lo[d] > reach_hi))
                return MX_ERR_BOX;
        }
    }
    // 1. ParallelCopy from the source valid boxes and their periodic images
    int nsh[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) nsh[d] = c.p.is_periodic[d] ? 1 : 0;
    for (size_t b = 0; b < mf.size(); ++b) {
        const IBox fb = fab_box(mf[b]);
        for (size_t f = 0; f < g.size(); ++f)
            for (int sz = -nsh[2]; sz <= nsh[2]; ++sz)
                for (int sy = -nsh[1]; sy <= nsh[1]; ++sy)
                    for (int sx = -nsh[0]; sx <= nsh[0]; ++sx) {
                        const int s[3] = {sx * (dom.hi[0] - dom.lo[0] + 1),
                                          sy * (dom.hi[1] - dom.lo[1] + 1),
                                          sz * (dom.hi[2] - dom.lo[2] + 1)};
                        if (b == f && s[0] == 0 && s[1] == 0 && s[2] == 0 &&
                            mf[b].p == src[f].p && dcomp == scomp)
                            continue;  // a FAB's own valid cells
                        IBox v = g[f];
                        for (int d = 0; d < 3; ++d) {
                            v.lo[d] += s[d];
                            v.hi[d] += s[d];
                        }
                        IBox r;
                        if (!intersect(fb, v, r)) continue;
                        const Arr& dst = mf[b];
                        const Arr& sa = src[f];
                        par_for(r, [&](int i, int j, int k) {
                            for (int n = 0; n < ncomp; ++n)
                                dst(i, j, k, dcomp + n) =
                                    sa(i - s[0], j - s[1], k - s[2], scomp + n);
                        });
                    }
    }
    // 2. fab_filcc, one direction at a time
    for (int d = 0; d < dim; ++d) {
        for (const Arr& q : mf) {
            const IBox fb = fab_box(q);
            if (inside(fb, gdom)) continue;
            const int ilo = dom.lo[d], ihi = dom.hi[d];
            const int is = fb.lo[d] > ilo ? fb.lo[d] : ilo;
            const int ie = fb.hi[d] < ihi ? fb.hi[d] : ihi;
            for (int side = 0; side < 2; ++side) {
                IBox r = fb;
                if (side == 0) {
                    if (fb.lo[d] >= ilo) continue;
                    r.hi[d] = ilo - 1;
                } else {
                    if (fb.hi[d] <= ihi) continue;
                    r.lo[d] = ihi + 1;
                }
                par_for(r, [&](int i, int j, int k) {
                    const int p[3] = {i, j, k};
                    const int x = p[d];
                    for (int n = 0; n < ncomp; ++n) {
                        const int cn = dcomp + n;
                        auto qv = [&](int y) -> double {
                            int t[3] = {i, j, k};
                            t[d] = y;
                            return q(t[0], t[1], t[2], cn);
                        };
                        const int bc = side ? bcs[n].hi[d] : bcs[n].lo[d];
                        double v;
                        if (side == 0) {
                            switch (bc) {
                                case MX_BC_FOEXTRAP: v = qv(ilo); break;
                                case MX_BC_HOEXTRAP:
                                    if (x < ilo - 1)
                                        v = qv(ilo);
                                    else if (ilo + 2 <= ie)
                                        v = 0.125 * (15. * qv(x + 1) - 10. * qv(x + 2) +
                                                     3. * qv(x + 3));
                                    else
                                        v = 0.5 * (3. * qv(x + 1) - qv(x + 2));
                                    break;
                                case MX_BC_REFLECT_EVEN: v = qv(2 * ilo - x - 1); break;
                                case MX_BC_REFLECT_ODD: v = -qv(2 * ilo - x - 1); break;
                                default: continue;
                            }
                        } else {
                            switch (bc) {
                                case MX_BC_FOEXTRAP: v = qv(ihi); break;
                                case MX_BC_HOEXTRAP:
                                    if (x > ihi + 1)
                                        v = qv(ihi);
                                    else if (ihi - 2 >= is)
                                        v = 0.125 * (15. * qv(x - 1) - 10. * qv(x - 2) +
                                                     3. * qv(x - 3));
                                    else
                                        v = 0.5 * (3. * qv(x - 1) - qv(x - 2));
                                    break;
                                case MX_BC_REFLECT_EVEN: v = qv(2 * ihi - x + 1); break;
                                case MX_BC_REFLECT_ODD: v = -qv(2 * ihi - x + 1); break;
                                default: continue;
                            }
                        }
                        q(i, j, k, cn) = v;
                    }
                });
            }
        }
    }
    // 3. FillExtBC: zero beyond ext_dir sides (Pi and velocity)
    for (int n = 0; n < ncomp; ++n) {
        const int cn = dcomp + n;
        if (!(is_vel || cn == c.p.comp_pi)) continue;
        for (int d = 0; d < dim; ++d)
            for (int side = 0; side < 2; ++side) {
                if ((side ? bcs[n].hi[d] : bcs[n].lo[d]) != MX_BC_EXT_DIR) continue;
                for (const Arr& q : mf) {
                    const IBox fb = fab_box(q);
                    if (inside(fb, gdom)) continue;
                    IBox r = fb;
                    if (side == 0) {
                        if (fb.lo[d] >= dom.lo[d]) continue;
                        r.hi[d] = dom.lo[d] - 1;
                    } else {
                        if (fb.hi[d] <= dom.hi[d]) continue;
                        r.lo[d] = dom.hi[d] + 1;
                    }
                    par_for(r, [&](int i, int j, int k) { q(i, j, k, cn) = 0.0; });
                }
            }
    }
    return MX_OK;
}

}  // namespace mxo
