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
#include <algorithm>
#include <array>
#include <cmath>
#include <utility>

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

// steps 2 and 3 on a list of FABs: fab_filcc one direction at a time, then
// FillExtBC, for the cells of each FAB beyond a side of `dom` in a
// non-periodic direction
void PhysFill(const Ctx& c, const IBox& dom, const std::vector<Arr>& fabs, int dcomp,
              int ncomp, const mx_bcrec* bcs, bool is_vel) {
    const int dim = c.dim();
    IBox gdom = dom;
    for (int d = 0; d < dim; ++d)
        if (c.p.is_periodic[d]) {
            gdom.lo[d] -= 1 << 24;
            gdom.hi[d] += 1 << 24;
        }
    // 2. fab_filcc, one direction at a time
    for (int d = 0; d < dim; ++d) {
        for (const Arr& q : fabs) {
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
                for (const Arr& q : fabs) {
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
}

// the checks made before anything is written: BCRec values, the FillExtBC
// abort (ext_dir on a scalar other than Pi with a FAB beyond that side) and
// that every cell a boundary rule reads lies in its FAB
int FillCheck(const Ctx& c, const IBox& dom, const std::vector<IBox>& fbs, int dcomp, int ncomp,
              const mx_bcrec* bcs, bool is_vel) {
    const int dim = c.dim();
    for (int n = 0; n < ncomp; ++n)
        for (int d = 0; d < dim; ++d)
            for (int side = 0; side < 2; ++side) {
                const int bc = side ? bcs[n].hi[d] : bcs[n].lo[d];
                if (bc < MX_BC_REFLECT_ODD || bc > MX_BC_HOEXTRAP) return MX_ERR_BC;
                if (bc != MX_BC_EXT_DIR || is_vel || dcomp + n == c.p.comp_pi) continue;
                for (const IBox& fb : fbs)
                    if (side ? fb.hi[d] > dom.hi[d] : fb.lo[d] < dom.lo[d]) return MX_ERR_BC;
            }
    IBox gdom = dom;
    for (int d = 0; d < dim; ++d)
        if (c.p.is_periodic[d]) {
            gdom.lo[d] -= 1 << 24;
            gdom.hi[d] += 1 << 24;
        }
    for (const IBox& fb : fbs) {
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
    return MX_OK;
}

int FillPatchSingleLevel(const Ctx& c, int lev, const std::vector<Arr>& mf,
                         const std::vector<Arr>& src, int scomp, int dcomp, int ncomp,
                         const mx_bcrec* bcs, bool is_vel) {
    const int dim = c.dim();
    const IBox dom = c.domain(lev);
    const std::vector<IBox>& g = c.grids[lev];
    std::vector<IBox> fbs;
    for (const Arr& a : mf) fbs.push_back(fab_box(a));
    int rc = FillCheck(c, dom, fbs, dcomp, ncomp, bcs, is_vel);
    if (rc) return rc;
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
    PhysFill(c, dom, mf, dcomp, ncomp, bcs, is_vel);
    return MX_OK;
}



// ---------------------------------------------------------------------------
// Maestro::FillPatch on a level lev > 0 (MaestroFillData.cpp:48-66):
// amrex::FillPatchTwoLevels with cell_cons_interp (CellConservativeLinear
// with linear limiting) and the physical-boundary functions above.
//
// TEST INFRASTRUCTURE ONLY.  FillPatchTwoLevels and CellConservativeLinear
// live in AMReX (AMReX_FillPatchUtil, AMReX_Interpolater, AMReX_MFInterp_C.H);
// this restates their published algorithm:
//   1. fine patches: per FAB of mf, its box intersected with the fine
//      domain grown in the periodic directions, minus every fine valid box
//      and its periodic images;
//   2. per fine patch, a coarse patch: the coarsened patch grown by one
//      coarse cell, filled by FillPatchSingleLevel on level lev - 1 (copy
//      from the coarse valid boxes and their periodic images, then the
//      physical boundary fill with the coarse domain);
//   3. the interpolation: on every coarse cell under the patch, per
//      direction d, the slope dc = (u(+1) - u(-1)) / 2, one-sided at a
//      domain side whose BCRec is ext_dir or hoextrap (the coarse patch
//      reaching that side: setBC), then the MC bound
//      s = sign(dc) min(min(|2 (u(+1) - u)|, |2 (u - u(-1))|) if
//      (u(+1) - u)(u - u(-1)) >= 0 else 0, |dc|), and a factor per
//      direction shared by all components, f_d = min(1, s / dc over the
//      components with dc != 0); the fine cell gets
//      u + xoff (dc_x f_x) + yoff (dc_y f_y) + zoff (dc_z f_z), with offsets
//      (i - ic r + 1/2) / r - 1/2;
//   4. the fine patches' values are copied into mf; then
//      FillPatchSingleLevel on level lev (copy from the fine valid boxes and
//      their periodic images, then the physical boundary fill).
// A coarse-patch cell inside the coarse domain that no coarse valid box (or
// periodic image) holds is an error (the levels are not properly nested
// for this ghost width): MX_ERR_BOX before anything is written.
namespace {

bool empty_b(const IBox& b) {
    for (int d = 0; d < 3; ++d)
        if (b.lo[d] > b.hi[d]) return true;
    return false;
}

// a \ b as disjoint boxes
void diff_box(const IBox& a, const IBox& b, std::vector<IBox>& out) {
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

// the level's valid boxes and their periodic images (shifts of -1, 0, +1
// domain lengths in the periodic directions)
struct Image {
    IBox first;                 // the shifted valid box
    std::array<int, 3> second;  // the shift
    size_t box;                 // the box's index
};

std::vector<Image> images(const Ctx& c, int lev) {
    const int dim = c.dim();
    const IBox dom = c.domain(lev);
    std::vector<Image> out;
    int nsh[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) nsh[d] = c.p.is_periodic[d] ? 1 : 0;
    for (size_t f = 0; f < c.grids[lev].size(); ++f)
        for (int sz = -nsh[2]; sz <= nsh[2]; ++sz)
            for (int sy = -nsh[1]; sy <= nsh[1]; ++sy)
                for (int sx = -nsh[0]; sx <= nsh[0]; ++sx) {
                    std::array<int, 3> s = {sx * (dom.hi[0] - dom.lo[0] + 1),
                                            sy * (dom.hi[1] - dom.lo[1] + 1),
                                            sz * (dom.hi[2] - dom.lo[2] + 1)};
                    IBox v = c.grids[lev][f];
                    for (int d = 0; d < 3; ++d) {
                        v.lo[d] += s[d];
                        v.hi[d] += s[d];
                    }
                    out.push_back({v, s, f});
                }
    return out;
}

int cfloor(int a, int r) { return a >= 0 ? a / r : -((-a + r - 1) / r); }

}  // namespace

// the fine patches of FillPatchTwoLevels (step 1) for FAB boxes fbs of level lev
std::vector<std::pair<int, IBox>> FinePatches(const Ctx& c, int lev, const std::vector<IBox>& fbs) {
    const int dim = c.dim();
    IBox fdg = c.domain(lev);
    for (int d = 0; d < dim; ++d)
        if (c.p.is_periodic[d]) {
            fdg.lo[d] -= 1 << 24;
            fdg.hi[d] += 1 << 24;
        }
    const auto imgs = images(c, lev);
    std::vector<std::pair<int, IBox>> out;
    for (size_t b = 0; b < fbs.size(); ++b) {
        IBox r;
        if (!intersect(fbs[b], fdg, r)) continue;
        std::vector<IBox> pieces{r};
        for (const auto& im : imgs) {
            std::vector<IBox> next;
            for (const IBox& q : pieces) diff_box(q, im.first, next);
            pieces.swap(next);
        }
        for (const IBox& q : pieces)
            if (!empty_b(q)) out.push_back({(int)b, q});
    }
    return out;
}

int FillPatchTwoLevels(const Ctx& c, int lev, const std::vector<Arr>& mf,
                       const std::vector<Arr>& crse, const std::vector<Arr>& fine, int scomp,
                       int dcomp, int ncomp, const mx_bcrec* bcs, bool is_vel) {
    const int dim = c.dim();
    int r[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) {
        const IBox a = c.domain(lev - 1), b = c.domain(lev);
        r[d] = (b.hi[d] - b.lo[d] + 1) / (a.hi[d] - a.lo[d] + 1);
    }
    const IBox cdom = c.domain(lev - 1);
    std::vector<IBox> fbs;
    for (const Arr& a : mf) fbs.push_back(fab_box(a));
    int rc = FillCheck(c, c.domain(lev), fbs, dcomp, ncomp, bcs, is_vel);
    if (rc) return rc;
    const auto patches = FinePatches(c, lev, fbs);
    // coarse patches and their coverage by the coarse level
    const auto cimgs = images(c, lev - 1);
    IBox cdg = cdom;
    for (int d = 0; d < dim; ++d)
        if (c.p.is_periodic[d]) {
            cdg.lo[d] -= 1 << 24;
            cdg.hi[d] += 1 << 24;
        }
    std::vector<IBox> cps;
    for (const auto& pt : patches) {
        IBox cp;
        for (int d = 0; d < 3; ++d) {
            cp.lo[d] = cfloor(pt.second.lo[d], r[d]) - (d < dim ? 1 : 0);
            cp.hi[d] = cfloor(pt.second.hi[d], r[d]) + (d < dim ? 1 : 0);
        }
        IBox in;
        if (intersect(cp, cdg, in)) {
            std::vector<IBox> pieces{in};
            for (const auto& im : cimgs) {
                std::vector<IBox> next;
                for (const IBox& q : pieces) diff_box(q, im.first, next);
                pieces.swap(next);
            }
            for (const IBox& q : pieces)
                if (!empty_b(q)) return MX_ERR_BOX;
        }
        cps.push_back(cp);
    }
    // the coarse patch holds the components at 0..ncomp-1, and that is the
    // component index its physical boundary functions see (FillExtBC's Pi
    // test included), as in the reference's FillPatchTwoLevels
    if ((rc = FillCheck(c, cdom, cps, 0, ncomp, bcs, is_vel))) return rc;
    // 2. coarse patch buffers: copy, then the coarse physical boundary fill
    std::vector<std::vector<double>> store(cps.size());
    std::vector<Arr> cpa(cps.size());
    for (size_t q = 0; q < cps.size(); ++q) {
        const IBox& b = cps[q];
        const long long nx = b.hi[0] - b.lo[0] + 1, ny = b.hi[1] - b.lo[1] + 1,
                        nz = b.hi[2] - b.lo[2] + 1;
        store[q].assign((size_t)(nx * ny * nz * ncomp), 0.0);
        Arr& a = cpa[q];
        a.p = store[q].data();
        a.js = nx;
        a.ks = nx * ny;
        a.ns = nx * ny * nz;
        for (int d = 0; d < 3; ++d) {
            a.b[d] = b.lo[d];
            a.e[d] = b.hi[d] + 1;
        }
        a.nc = ncomp;
        for (const auto& im : cimgs) {
            IBox x;
            if (!intersect(b, im.first, x)) continue;
            const Arr& sa = crse[im.box];
            const auto& s = im.second;
            par_for(x, [&](int i, int j, int k) {
                for (int n = 0; n < ncomp; ++n)
                    a(i, j, k, n) = sa(i - s[0], j - s[1], k - s[2], scomp + n);
            });
        }
    }
    std::vector<mx_bcrec> bc0(bcs, bcs + ncomp);
    PhysFill(c, cdom, cpa, 0, ncomp, bc0.data(), is_vel);
    // 3. interpolation into the fine patches
    for (size_t q = 0; q < patches.size(); ++q) {
        const Arr& u = cpa[q];
        const IBox& cp = cps[q];
        const Arr& dst = mf[patches[q].first];
        // setBC: the domain's BCRec on a side the coarse patch reaches
        std::vector<int> onesided_lo(ncomp * 3, 0), onesided_hi(ncomp * 3, 0);
        for (int n = 0; n < ncomp; ++n)
            for (int d = 0; d < dim; ++d) {
                const int blo = cp.lo[d] <= cdom.lo[d] ? bcs[n].lo[d] : MX_BC_INT_DIR;
                const int bhi = cp.hi[d] >= cdom.hi[d] ? bcs[n].hi[d] : MX_BC_INT_DIR;
                onesided_lo[n * 3 + d] = blo == MX_BC_EXT_DIR || blo == MX_BC_HOEXTRAP;
                onesided_hi[n * 3 + d] = bhi == MX_BC_EXT_DIR || bhi == MX_BC_HOEXTRAP;
            }
        auto central = [&](int d, int n, int ic, int jc, int kc) {
            const int e0 = d == 0, e1 = d == 1, e2 = d == 2;
            auto U = [&](int o) { return u(ic + o * e0, jc + o * e1, kc + o * e2, n); };
            const int x = d == 0 ? ic : (d == 1 ? jc : kc);
            double dc = 0.5 * (U(1) - U(-1));
            if (x == cdom.lo[d] && onesided_lo[n * 3 + d]) {
                if (x + 2 <= cp.hi[d])
                    dc = -16. / 15. * U(-1) + 0.5 * U(0) + 2. / 3. * U(1) - 0.1 * U(2);
                else
                    dc = 0.25 * (U(1) + 5. * U(0) - 6. * U(-1));
            }
            if (x == cdom.hi[d] && onesided_hi[n * 3 + d]) {
                if (x - 2 >= cp.lo[d])
                    dc = 16. / 15. * U(1) - 0.5 * U(0) - 2. / 3. * U(-1) + 0.1 * U(-2);
                else
                    dc = -0.25 * (U(-1) + 5. * U(0) - 6. * U(1));
            }
            return dc;
        };
        par_for(patches[q].second, [&](int i, int j, int k) {
            const int ic = cfloor(i, r[0]), jc = cfloor(j, r[1]), kc = cfloor(k, r[2]);
            double sf[3] = {1.0, 1.0, 1.0};
            for (int n = 0; n < ncomp; ++n)
                for (int d = 0; d < dim; ++d) {
                    const int e0 = d == 0, e1 = d == 1, e2 = d == 2;
                    const double u0 = u(ic, jc, kc, n);
                    const double up = u(ic + e0, jc + e1, kc + e2, n);
                    const double um = u(ic - e0, jc - e1, kc - e2, n);
                    const double dc = central(d, n, ic, jc, kc);
                    const double df = 2.0 * (up - u0);
                    const double db = 2.0 * (u0 - um);
                    double s = (df * db >= 0.0) ? std::min(std::abs(df), std::abs(db)) : 0.0;
                    s = std::copysign(1.0, dc) * std::min(s, std::abs(dc));
                    if (dc != 0.0) sf[d] = std::min(sf[d], s / dc);
                }
            const double off[3] = {(i - ic * r[0] + 0.5) / r[0] - 0.5,
                                   (j - jc * r[1] + 0.5) / r[1] - 0.5,
                                   (k - kc * r[2] + 0.5) / r[2] - 0.5};
            for (int n = 0; n < ncomp; ++n) {
                double v = u(ic, jc, kc, n);
                for (int d = 0; d < dim; ++d) v = v + off[d] * (central(d, n, ic, jc, kc) * sf[d]);
                dst(i, j, k, dcomp + n) = v;
            }
        });
    }
    // 4. the fine level's own valid data and physical boundaries
    return FillPatchSingleLevel(c, lev, mf, fine, scomp, dcomp, ncomp, bcs, is_vel);
}

}  // namespace mxo
