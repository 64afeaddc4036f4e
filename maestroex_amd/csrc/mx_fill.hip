// mx_fill.hip -- Maestro::FillPatch at level 0 (MaestroFillData.cpp:24-47):
// amrex::FillPatchSingleLevel with Maestro's ScalarFill / VelFill
// (MaestroBCFill.cpp:8-28) and FillExtBC (:30-171), on the device.
//
//   1. k_fill_copy: every cell of a destination FAB's grown box that lies in
//      a source valid box or one of its periodic images takes the source
//      value.  The host intersects the boxes; each intersection is a task.
//      Valid boxes are disjoint, so every destination cell has one source.
//   2. k_fill_phys<D, false>, D = x, y, z in turn: fab_filcc's rule for
//      direction D (foextrap, hoextrap, reflect_even/odd) on the cells of
//      every FAB beyond the domain in D.  A pass reads only cells inside the
//      domain in D, which no launch of the pass writes.  One pass per
//      direction fixes the order at corners outside the domain in two
//      non-periodic directions: the later direction's rule, applied to the
//      values the earlier pass filled.
//   3. k_fill_phys<D, true>: FillExtBC, zero beyond ext_dir sides for Pi and
//      velocity, after every rule has run.
// HBM-bound: one read and one write per ghost cell.
#include "mx_internal.H"

using namespace mxd;
using namespace mxh;

namespace {

constexpr int NTF = 256;

struct CopyTask {
    int lo[3], hi[3];  // destination cells
    int sh[3];         // periodic shift: source cell = destination cell - sh
    int blk0;          // the task's first block
    DA dst, src;
};

struct CopyArgs {
    int scomp, dcomp;
};

struct BcTask {
    int lo[3], hi[3];  // cells beyond the domain on one side in direction D
    int side;          // 0: below domlo, 1: above domhi
    int blk0;
    int is, ie;  // the FAB's extent in D clipped to the domain
    DA q;
};

struct BcArgs {
    int dcomp;
    int dlo, dhi;  // domain in D
    int comp_pi;   // component index of Pi
    int is_vel;
    int pad_ = 0;
};

__global__ void __launch_bounds__(NTF) k_fill_copy(const CopyTask* __restrict__ tasks,
                                                   const int* __restrict__ blk, CopyArgs A) {
    const CopyTask T = tasks[blk[blockIdx.x]];
    int i, j, k;
    {
        const long long idx = (long long)(blockIdx.x - T.blk0) * NTF + threadIdx.x;
        if (!region_point(idx, T.lo, T.hi, i, j, k)) return;
    }
    const int n = (int)blockIdx.z;
    T.dst(i, j, k, A.dcomp + n) = T.src(i - T.sh[0], j - T.sh[1], k - T.sh[2], A.scomp + n);
}

template <int D, bool ZERO>
__global__ void __launch_bounds__(NTF) k_fill_phys(const BcTask* __restrict__ tasks,
                                                   const int* __restrict__ blk,
                                                   const mx_bcrec* __restrict__ bcs, BcArgs A) {
    const BcTask T = tasks[blk[blockIdx.x]];
    int i, j, k;
    {
        const long long idx = (long long)(blockIdx.x - T.blk0) * NTF + threadIdx.x;
        if (!region_point(idx, T.lo, T.hi, i, j, k)) return;
    }
    const int n = (int)blockIdx.z;
    const int cn = A.dcomp + n;
    const int bc = T.side ? bcs[n].hi[D] : bcs[n].lo[D];
    double* q = T.q.p + T.q.off(i, j, k, cn);
    if (ZERO) {
        if (bc == MX_BC_EXT_DIR && (A.is_vel || cn == A.comp_pi)) *q = 0.0;
        return;
    }
    const long long sd = D == 0 ? 1 : (D == 1 ? T.q.js : T.q.ks);
    const int x = D == 0 ? i : (D == 1 ? j : k);
    auto at = [&](int y) -> double { return q[(long long)(y - x) * sd]; };
    double v;
    if (T.side == 0) {
        const int ilo = A.dlo;
        switch (bc) {
            case MX_BC_FOEXTRAP: v = at(ilo); break;
            case MX_BC_HOEXTRAP:
                if (x < ilo - 1)
                    v = at(ilo);
                else if (ilo + 2 <= T.ie)
                    v = 0.125 * (15. * at(x + 1) - 10. * at(x + 2) + 3. * at(x + 3));
                else
                    v = 0.5 * (3. * at(x + 1) - at(x + 2));
                break;
            case MX_BC_REFLECT_EVEN: v = at(2 * ilo - x - 1); break;
            case MX_BC_REFLECT_ODD: v = -at(2 * ilo - x - 1); break;
            default: return;
        }
    } else {
        const int ihi = A.dhi;
        switch (bc) {
            case MX_BC_FOEXTRAP: v = at(ihi); break;
            case MX_BC_HOEXTRAP:
                if (x > ihi + 1)
                    v = at(ihi);
                else if (ihi - 2 >= T.is)
                    v = 0.125 * (15. * at(x - 1) - 10. * at(x - 2) + 3. * at(x - 3));
                else
                    v = 0.5 * (3. * at(x - 1) - at(x - 2));
                break;
            case MX_BC_REFLECT_EVEN: v = at(2 * ihi - x + 1); break;
            case MX_BC_REFLECT_ODD: v = -at(2 * ihi - x + 1); break;
            default: return;
        }
    }
    *q = v;
}

bool inside(const HBox& a, const HBox& b) {
    for (int d = 0; d < 3; ++d)
        if (a.lo[d] < b.lo[d] || a.hi[d] > b.hi[d]) return false;
    return true;
}

bool intersect(const HBox& a, const HBox& b, HBox& o) {
    for (int d = 0; d < 3; ++d) {
        o.lo[d] = std::max(a.lo[d], b.lo[d]);
        o.hi[d] = std::min(a.hi[d], b.hi[d]);
        if (o.lo[d] > o.hi[d]) return false;
    }
    return true;
}

HBox fab_box(const mx_fab& f) {
    HBox b;
    for (int d = 0; d < 3; ++d) {
        b.lo[d] = f.begin[d];
        b.hi[d] = f.end[d] - 1;
    }
    return b;
}

// block-to-task table: every task gets ceil(points / NTF) consecutive blocks
template <class T>
int flatten(std::vector<T>& tasks, std::vector<int>& blk) {
    blk.clear();
    for (size_t q = 0; q < tasks.size(); ++q) {
        HBox r;
        for (int d = 0; d < 3; ++d) {
            r.lo[d] = tasks[q].lo[d];
            r.hi[d] = tasks[q].hi[d];
        }
        const long long nb = (npts(r) + NTF - 1) / NTF;
        if ((long long)blk.size() + nb > 0x7fffffffLL) return MX_ERR_UNSUPPORTED;
        tasks[q].blk0 = (int)blk.size();
        blk.insert(blk.end(), (size_t)nb, (int)q);
    }
    return MX_OK;
}

template <int D>
int launch_phys(mx_adv_ctx* c, const std::vector<BcTask>& tasks, const std::vector<int>& blk,
                const std::vector<mx_bcrec>& bcs, const BcArgs& A, int ncomp, bool zero) {
    if (tasks.empty()) return MX_OK;
    Table tab(c);
    const size_t ot = tab.add(tasks);
    const size_t ob = tab.add(blk);
    const size_t oc = tab.add(bcs);
    int rc = tab.upload();
    if (rc) return rc;
    const dim3 grid((unsigned)blk.size(), 1, (unsigned)ncomp);
    if (zero) {
        MX_LAUNCH(c, "k_fill_phys", (k_fill_phys<D, true>), grid, dim3(NTF),
                  tab.dev<BcTask>(ot), tab.dev<int>(ob), tab.dev<mx_bcrec>(oc), A);
    } else {
        MX_LAUNCH(c, "k_fill_phys", (k_fill_phys<D, false>), grid, dim3(NTF),
                  tab.dev<BcTask>(ot), tab.dev<int>(ob), tab.dev<mx_bcrec>(oc), A);
    }
    return MX_OK;
}

}  // namespace

extern "C" {

int mx_adv_fill_patch(mx_adv_ctx* c, int lev, mx_multifab* mf, const mx_multifab* src,
                      int srccomp, int destcomp, int ncomp, const mx_bcrec* bcs,
                      int variable_type) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    if (lev != 0) return MX_ERR_UNSUPPORTED;  // FillPatchTwoLevels (coarse interpolation)
    const auto& grids = c->grids[lev];
    const int nb = (int)grids.size();
    const int dim = c->p.dim;
    if (!mf || !src || !bcs || mf->nfab != nb || src->nfab != nb || !mf->fab || !src->fab ||
        srccomp < 0 || destcomp < 0 || ncomp < 1 || ncomp > 65535)
        return MX_ERR_ARG;
    // a MultiFab copied onto itself with shifted, overlapping component
    // ranges would read components the same launch writes
    if (mf->fab == src->fab || (nb > 0 && mf->fab[0].p == src->fab[0].p))
        if (srccomp != destcomp && srccomp < destcomp + ncomp && destcomp < srccomp + ncomp)
            return MX_ERR_ARG;
    const bool is_vel = variable_type == 1;
    HBox dom;
    for (int d = 0; d < 3; ++d) {
        dom.lo[d] = d < dim ? c->p.domain_lo[lev][d] : 0;
        dom.hi[d] = d < dim ? c->p.domain_hi[lev][d] : 0;
    }
    // FABs: the valid box grown by one ghost width, the same for all
    int ng[3] = {0, 0, 0};
    for (int b = 0; b < nb; ++b) {
        const mx_fab& m = mf->fab[b];
        const mx_fab& s = src->fab[b];
        const HBox g = hbox(grids[b]);
        if (!m.p || !s.p || m.ncomp < destcomp + ncomp || s.ncomp < srccomp + ncomp)
            return MX_ERR_ARG;
        for (int d = 0; d < 3; ++d) {
            const int w = g.lo[d] - m.begin[d];
            if (w < 0 || m.end[d] - 1 - g.hi[d] != w || s.begin[d] > g.lo[d] ||
                s.end[d] - 1 < g.hi[d])
                return MX_ERR_BOX;
            if (b == 0) ng[d] = w;
            if (w != ng[d]) return MX_ERR_BOX;
        }
    }
    // BCRec values; a scalar other than Pi with ext_dir is the reference's abort
    for (int n = 0; n < ncomp; ++n)
        for (int d = 0; d < dim; ++d)
            for (int side = 0; side < 2; ++side) {
                const int bc = side ? bcs[n].hi[d] : bcs[n].lo[d];
                if (bc < MX_BC_REFLECT_ODD || bc > MX_BC_HOEXTRAP) return MX_ERR_BC;
                if (bc != MX_BC_EXT_DIR || is_vel || destcomp + n == c->p.comp_pi) continue;
                for (int b = 0; b < nb; ++b) {
                    const HBox fb = fab_box(mf->fab[b]);
                    if (side ? fb.hi[d] > dom.hi[d] : fb.lo[d] < dom.lo[d]) return MX_ERR_BC;
                }
            }
    HBox gdom = dom;
    for (int d = 0; d < dim; ++d)
        if (c->p.is_periodic[d]) {
            gdom.lo[d] -= ng[d];
            gdom.hi[d] += ng[d];
        }
    // every cell a boundary rule reads lies in the FAB
    for (int b = 0; b < nb; ++b) {
        const HBox fb = fab_box(mf->fab[b]);
        if (inside(fb, gdom)) continue;
        for (int d = 0; d < dim; ++d) {
            const int reach_lo = 2 * dom.lo[d] - fb.lo[d] - 1;
            const int reach_hi = 2 * dom.hi[d] - fb.hi[d] + 1;
            if (fb.lo[d] < dom.lo[d] && (fb.hi[d] < dom.lo[d] + 1 || fb.hi[d] < reach_lo))
                return MX_ERR_BOX;
            if (fb.hi[d] > dom.hi[d] && (fb.lo[d] > dom.hi[d] - 1 || fb.lo[d] > reach_hi))
                return MX_ERR_BOX;
        }
    }

    // 1. ParallelCopy from the source valid boxes and their periodic images;
    //    the box intersections are planned once per (level, ghost widths)
    {
        const auto key = std::make_tuple(lev, ng[0], ng[1], ng[2]);
        auto it = c->fill_plans.find(key);
        if (it == c->fill_plans.end()) {
            int nsh[3] = {0, 0, 0};
            for (int d = 0; d < dim; ++d) nsh[d] = c->p.is_periodic[d] ? 1 : 0;
            std::vector<mx_adv_ctx::FillCopy> plan;
            for (int b = 0; b < nb; ++b) {
                HBox gb = hbox(grids[b]);  // the FAB's box: grown by ng (checked above)
                for (int d = 0; d < 3; ++d) {
                    gb.lo[d] -= ng[d];
                    gb.hi[d] += ng[d];
                }
                for (int f = 0; f < nb; ++f)
                    for (int sz = -nsh[2]; sz <= nsh[2]; ++sz)
                        for (int sy = -nsh[1]; sy <= nsh[1]; ++sy)
                            for (int sx = -nsh[0]; sx <= nsh[0]; ++sx) {
                                const int s[3] = {sx * (dom.hi[0] - dom.lo[0] + 1),
                                                  sy * (dom.hi[1] - dom.lo[1] + 1),
                                                  sz * (dom.hi[2] - dom.lo[2] + 1)};
                                HBox v = hbox(grids[f]);
                                for (int d = 0; d < 3; ++d) {
                                    v.lo[d] += s[d];
                                    v.hi[d] += s[d];
                                }
                                HBox r;
                                if (!intersect(gb, v, r)) continue;
                                mx_adv_ctx::FillCopy q;
                                q.b = b;
                                q.f = f;
                                for (int d = 0; d < 3; ++d) {
                                    q.sh[d] = s[d];
                                    q.lo[d] = r.lo[d];
                                    q.hi[d] = r.hi[d];
                                }
                                plan.push_back(q);
                            }
            }
            it = c->fill_plans.emplace(key, std::move(plan)).first;
        }
        std::vector<CopyTask> tasks;
        tasks.reserve(it->second.size());
        for (const auto& q : it->second) {
            if (q.b == q.f && !q.sh[0] && !q.sh[1] && !q.sh[2] &&
                mf->fab[q.b].p == src->fab[q.f].p && destcomp == srccomp)
                continue;  // a FAB's own valid cells
            CopyTask T;
            for (int d = 0; d < 3; ++d) {
                T.lo[d] = q.lo[d];
                T.hi[d] = q.hi[d];
                T.sh[d] = q.sh[d];
            }
            T.blk0 = 0;
            T.dst = to_da(mf->fab[q.b]);
            T.src = to_da(src->fab[q.f]);
            tasks.push_back(T);
        }
        std::vector<int> blk;
        if ((rc = flatten(tasks, blk))) return rc;
        if (!tasks.empty()) {
            Table tab(c);
            const size_t ot = tab.add(tasks);
            const size_t ob = tab.add(blk);
            if ((rc = tab.upload())) return rc;
            CopyArgs A{srccomp, destcomp};
            MX_LAUNCH(c, "k_fill_copy", k_fill_copy, dim3((unsigned)blk.size(), 1, (unsigned)ncomp),
                      dim3(NTF), tab.dev<CopyTask>(ot), tab.dev<int>(ob), A);
        }
    }

    // 2. fab_filcc one direction at a time, then 3. FillExtBC
    const std::vector<mx_bcrec> bcv(bcs, bcs + ncomp);
    // the sides with work: a boundary rule (pass 2) or a zeroed component
    // (pass 3) for some component; int_dir sides keep their cells
    auto side_has = [&](int d, int side, bool zero) {
        for (int n = 0; n < ncomp; ++n) {
            const int bc = side ? bcs[n].hi[d] : bcs[n].lo[d];
            if (zero ? (bc == MX_BC_EXT_DIR && (is_vel || destcomp + n == c->p.comp_pi))
                     : (bc != MX_BC_INT_DIR && bc != MX_BC_EXT_DIR))
                return true;
        }
        return false;
    };
    for (int zero = 0; zero < 2; ++zero) {
        std::vector<BcTask> per_dir[3];
        std::vector<int> blk_dir[3];
        for (int d = 0; d < dim; ++d) {
            for (int b = 0; b < nb; ++b) {
                const HBox fb = fab_box(mf->fab[b]);
                if (inside(fb, gdom)) continue;
                for (int side = 0; side < 2; ++side) {
                    if (!side_has(d, side, zero != 0)) continue;
                    HBox r = fb;
                    if (side == 0) {
                        if (fb.lo[d] >= dom.lo[d]) continue;
                        r.hi[d] = dom.lo[d] - 1;
                    } else {
                        if (fb.hi[d] <= dom.hi[d]) continue;
                        r.lo[d] = dom.hi[d] + 1;
                    }
                    BcTask T;
                    for (int q = 0; q < 3; ++q) {
                        T.lo[q] = r.lo[q];
                        T.hi[q] = r.hi[q];
                    }
                    T.side = side;
                    T.blk0 = 0;
                    T.is = std::max(fb.lo[d], dom.lo[d]);
                    T.ie = std::min(fb.hi[d], dom.hi[d]);
                    T.q = to_da(mf->fab[b]);
                    per_dir[d].push_back(T);
                }
            }
            if ((rc = flatten(per_dir[d], blk_dir[d]))) return rc;
        }
        for (int d = 0; d < dim; ++d) {
            BcArgs A;
            A.dcomp = destcomp;
            A.dlo = dom.lo[d];
            A.dhi = dom.hi[d];
            A.comp_pi = c->p.comp_pi;
            A.is_vel = is_vel ? 1 : 0;
            if (d == 0)
                rc = launch_phys<0>(c, per_dir[d], blk_dir[d], bcv, A, ncomp, zero != 0);
            else if (d == 1)
                rc = launch_phys<1>(c, per_dir[d], blk_dir[d], bcv, A, ncomp, zero != 0);
            else
                rc = launch_phys<2>(c, per_dir[d], blk_dir[d], bcv, A, ncomp, zero != 0);
            if (rc) return rc;
        }
    }
    return MX_OK;
}

}  // extern "C"
