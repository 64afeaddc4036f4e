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

// FillPatchTwoLevels' interpolation (CellConservativeLinear with linear
// limiting): a task is one fine patch with its coarse patch buffer u (the
// patch coarsened and grown by one coarse cell)
struct InterpTask {
    int lo[3], hi[3];    // the fine patch's cells
    int clo[3], chi[3];  // the coarse patch
    int blk0;
    unsigned one_lo[3], one_hi[3];  // per direction, bit n: one-sided slope
                                    // of component n at the domain side
    DA u, dst;
};

struct InterpArgs {
    int r[3];
    int dlo[3], dhi[3];  // the coarse domain
    int dim, ncomp, dcomp;
};

__device__ __forceinline__ int cfloor_d(int a, int r) { return a >= 0 ? a / r : -((-a + r - 1) / r); }

// the central slope of component n in direction d at coarse cell c, one-sided
// at a domain side where the coarse patch reaches it with ext_dir / hoextrap
__device__ __forceinline__ double interp_dc(const InterpTask& T, const InterpArgs& A, int d, int n,
                                            int ic, int jc, int kc) {
    const int e0 = d == 0, e1 = d == 1, e2 = d == 2;
    auto U = [&](int o) { return T.u(ic + o * e0, jc + o * e1, kc + o * e2, n); };
    const int x = d == 0 ? ic : (d == 1 ? jc : kc);
    double dc = 0.5 * (U(1) - U(-1));
    if (x == A.dlo[d] && ((T.one_lo[d] >> n) & 1u)) {
        if (x + 2 <= T.chi[d])
            dc = -16. / 15. * U(-1) + 0.5 * U(0) + 2. / 3. * U(1) - 0.1 * U(2);
        else
            dc = 0.25 * (U(1) + 5. * U(0) - 6. * U(-1));
    }
    if (x == A.dhi[d] && ((T.one_hi[d] >> n) & 1u)) {
        if (x - 2 >= T.clo[d])
            dc = 16. / 15. * U(1) - 0.5 * U(0) - 2. / 3. * U(-1) + 0.1 * U(-2);
        else
            dc = -0.25 * (U(-1) + 5. * U(0) - 6. * U(1));
    }
    return dc;
}

__global__ void __launch_bounds__(NTF) k_fill_interp(const InterpTask* __restrict__ tasks,
                                                     const int* __restrict__ blk, InterpArgs A) {
    const InterpTask T = tasks[blk[blockIdx.x]];
    int i, j, k;
    {
        const long long idx = (long long)(blockIdx.x - T.blk0) * NTF + threadIdx.x;
        if (!region_point(idx, T.lo, T.hi, i, j, k)) return;
    }
    const int ic = cfloor_d(i, A.r[0]), jc = cfloor_d(j, A.r[1]), kc = cfloor_d(k, A.r[2]);
    // the limiting factor of each direction, shared by the components
    double sf[3] = {1.0, 1.0, 1.0};
    for (int n = 0; n < A.ncomp; ++n)
        for (int d = 0; d < A.dim; ++d) {
            const int e0 = d == 0, e1 = d == 1, e2 = d == 2;
            const double u0 = T.u(ic, jc, kc, n);
            const double up = T.u(ic + e0, jc + e1, kc + e2, n);
            const double um = T.u(ic - e0, jc - e1, kc - e2, n);
            const double dc = interp_dc(T, A, d, n, ic, jc, kc);
            const double df = 2.0 * (up - u0);
            const double db = 2.0 * (u0 - um);
            double s = (df * db >= 0.0) ? dmin(dabs(df), dabs(db)) : 0.0;
            s = copysign(1.0, dc) * dmin(s, dabs(dc));
            if (dc != 0.0) sf[d] = dmin(sf[d], s / dc);
        }
    const double off[3] = {(i - ic * A.r[0] + 0.5) / A.r[0] - 0.5,
                           (j - jc * A.r[1] + 0.5) / A.r[1] - 0.5,
                           (k - kc * A.r[2] + 0.5) / A.r[2] - 0.5};
    for (int n = 0; n < A.ncomp; ++n) {
        double v = T.u(ic, jc, kc, n);
        for (int d = 0; d < A.dim; ++d) v = v + off[d] * (interp_dc(T, A, d, n, ic, jc, kc) * sf[d]);
        T.dst(i, j, k, A.dcomp + n) = v;
    }
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


// fab_filcc one direction at a time, then FillExtBC, on the cells of each
// FAB beyond a side of `dom` in a non-periodic direction; `dcomp` is the
// component index the boundary functions see (FillExtBC's Pi test)
int phys_fill(mx_adv_ctx* c, const HBox& dom, const std::vector<std::pair<DA, HBox>>& fabs,
              int dcomp, int ncomp, const mx_bcrec* bcs, bool is_vel) {
    const int dim = c->p.dim;
    HBox gdom = dom;
    for (int d = 0; d < dim; ++d)
        if (c->p.is_periodic[d]) {
            gdom.lo[d] -= 1 << 24;
            gdom.hi[d] += 1 << 24;
        }
    const std::vector<mx_bcrec> bcv(bcs, bcs + ncomp);
    // the sides with work: a boundary rule (pass 2) or a zeroed component
    // (pass 3) for some component; int_dir sides keep their cells
    auto side_has = [&](int d, int side, bool zero) {
        for (int n = 0; n < ncomp; ++n) {
            const int bc = side ? bcs[n].hi[d] : bcs[n].lo[d];
            if (zero ? (bc == MX_BC_EXT_DIR && (is_vel || dcomp + n == c->p.comp_pi))
                     : (bc != MX_BC_INT_DIR && bc != MX_BC_EXT_DIR))
                return true;
        }
        return false;
    };
    int rc;
    for (int zero = 0; zero < 2; ++zero) {
        std::vector<BcTask> per_dir[3];
        std::vector<int> blk_dir[3];
        for (int d = 0; d < dim; ++d) {
            for (const auto& fab : fabs) {
                const HBox& fb = fab.second;
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
                    T.q = fab.first;
                    per_dir[d].push_back(T);
                }
            }
            if ((rc = flatten(per_dir[d], blk_dir[d]))) return rc;
        }
        for (int d = 0; d < dim; ++d) {
            BcArgs A;
            A.dcomp = dcomp;
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

// the checks made before anything is written, on FAB boxes `fbs` and the
// domain `dom`: BCRec values, the FillExtBC abort (ext_dir on a scalar
// other than Pi with a FAB beyond that side; `dcomp` as in phys_fill) and
// that every cell a boundary rule reads lies in its FAB
int phys_check(const mx_adv_ctx* c, const HBox& dom, const std::vector<HBox>& fbs, int dcomp,
               int ncomp, const mx_bcrec* bcs, bool is_vel) {
    const int dim = c->p.dim;
    for (int n = 0; n < ncomp; ++n)
        for (int d = 0; d < dim; ++d)
            for (int side = 0; side < 2; ++side) {
                const int bc = side ? bcs[n].hi[d] : bcs[n].lo[d];
                if (bc < MX_BC_REFLECT_ODD || bc > MX_BC_HOEXTRAP) return MX_ERR_BC;
                if (bc != MX_BC_EXT_DIR || is_vel || dcomp + n == c->p.comp_pi) continue;
                for (const HBox& fb : fbs)
                    if (side ? fb.hi[d] > dom.hi[d] : fb.lo[d] < dom.lo[d]) return MX_ERR_BC;
            }
    HBox gdom = dom;
    for (int d = 0; d < dim; ++d)
        if (c->p.is_periodic[d]) {
            gdom.lo[d] -= 1 << 24;
            gdom.hi[d] += 1 << 24;
        }
    for (const HBox& fb : fbs) {
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
    return MX_OK;
}

HBox level_domain(const mx_adv_ctx* c, int lev) {
    HBox dom;
    for (int d = 0; d < 3; ++d) {
        dom.lo[d] = d < c->p.dim ? c->p.domain_lo[lev][d] : 0;
        dom.hi[d] = d < c->p.dim ? c->p.domain_hi[lev][d] : 0;
    }
    return dom;
}

// argument checks of FillPatchSingleLevel on level lev (mf's FABs: the
// valid boxes grown by one ghost width ng, the same for all)
int single_check(mx_adv_ctx* c, int lev, const mx_multifab* mf, const mx_multifab* src,
                 int srccomp, int destcomp, int ncomp, const mx_bcrec* bcs, bool is_vel,
                 int ng[3]) {
    const auto& grids = c->grids[lev];
    const int nb = (int)grids.size();
    if (!mf || !src || !bcs || mf->nfab != nb || src->nfab != nb || !mf->fab || !src->fab ||
        srccomp < 0 || destcomp < 0 || ncomp < 1 || ncomp > 65535)
        return MX_ERR_ARG;
    // a MultiFab copied onto itself with shifted, overlapping component
    // ranges would read components the same launch writes
    if (mf->fab == src->fab || (nb > 0 && mf->fab[0].p == src->fab[0].p))
        if (srccomp != destcomp && srccomp < destcomp + ncomp && destcomp < srccomp + ncomp)
            return MX_ERR_ARG;
    for (int d = 0; d < 3; ++d) ng[d] = 0;
    std::vector<HBox> fbs;
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
        fbs.push_back(fab_box(m));
    }
    return phys_check(c, level_domain(c, lev), fbs, destcomp, ncomp, bcs, is_vel);
}

// FillPatchSingleLevel on level lev after single_check: the copy from the
// source valid boxes and their periodic images, then the physical boundaries
int single_run(mx_adv_ctx* c, int lev, mx_multifab* mf, const mx_multifab* src, int srccomp,
               int destcomp, int ncomp, const mx_bcrec* bcs, bool is_vel, const int ng[3]) {
    const auto& grids = c->grids[lev];
    const int nb = (int)grids.size();
    const int dim = c->p.dim;
    const HBox dom = level_domain(c, lev);
    int rc;
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

    std::vector<std::pair<DA, HBox>> fabs;
    for (int b = 0; b < nb; ++b) fabs.push_back({to_da(mf->fab[b]), fab_box(mf->fab[b])});
    return phys_fill(c, dom, fabs, destcomp, ncomp, bcs, is_vel);
}


// a \ b as disjoint boxes (the order fixes the fine patches' shapes)
void diff_box(const HBox& a, const HBox& b, std::vector<HBox>& out) {
    HBox in;
    if (!intersect(a, b, in)) {
        out.push_back(a);
        return;
    }
    HBox rest = a;
    for (int d = 0; d < 3; ++d) {
        if (rest.lo[d] < in.lo[d]) {
            HBox s = rest;
            s.hi[d] = in.lo[d] - 1;
            out.push_back(s);
            rest.lo[d] = in.lo[d];
        }
        if (rest.hi[d] > in.hi[d]) {
            HBox s = rest;
            s.lo[d] = in.hi[d] + 1;
            out.push_back(s);
            rest.hi[d] = in.hi[d];
        }
    }
}

struct Image {
    HBox box;  // a valid box, shifted
    int sh[3];
    int src;   // the box's index
};

// the level's valid boxes and their periodic images, box by box, shifts
// z outermost
std::vector<Image> level_images(const mx_adv_ctx* c, int lev) {
    const int dim = c->p.dim;
    const HBox dom = level_domain(c, lev);
    int nsh[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) nsh[d] = c->p.is_periodic[d] ? 1 : 0;
    std::vector<Image> out;
    for (size_t f = 0; f < c->grids[lev].size(); ++f)
        for (int sz = -nsh[2]; sz <= nsh[2]; ++sz)
            for (int sy = -nsh[1]; sy <= nsh[1]; ++sy)
                for (int sx = -nsh[0]; sx <= nsh[0]; ++sx) {
                    Image im;
                    im.sh[0] = sx * (dom.hi[0] - dom.lo[0] + 1);
                    im.sh[1] = sy * (dom.hi[1] - dom.lo[1] + 1);
                    im.sh[2] = sz * (dom.hi[2] - dom.lo[2] + 1);
                    im.box = hbox(c->grids[lev][f]);
                    for (int d = 0; d < 3; ++d) {
                        im.box.lo[d] += im.sh[d];
                        im.box.hi[d] += im.sh[d];
                    }
                    im.src = (int)f;
                    out.push_back(im);
                }
    return out;
}

// the part of `r` no image holds
std::vector<HBox> uncovered(const HBox& r, const std::vector<Image>& imgs) {
    std::vector<HBox> pieces{r};
    for (const Image& im : imgs) {
        std::vector<HBox> next;
        for (const HBox& q : pieces) diff_box(q, im.box, next);
        pieces.swap(next);
    }
    std::vector<HBox> out;
    for (const HBox& q : pieces) {
        bool empty = false;
        for (int d = 0; d < 3; ++d)
            if (q.lo[d] > q.hi[d]) empty = true;
        if (!empty) out.push_back(q);
    }
    return out;
}

int floordiv_h(int a, int r) { return a >= 0 ? a / r : -((-a + r - 1) / r); }

HBox periodic_grown(const mx_adv_ctx* c, HBox dom) {
    for (int d = 0; d < c->p.dim; ++d)
        if (c->p.is_periodic[d]) {
            dom.lo[d] -= 1 << 24;
            dom.hi[d] += 1 << 24;
        }
    return dom;
}

}  // namespace

extern "C" {

int mx_adv_fill_patch(mx_adv_ctx* c, int lev, mx_multifab* mf, const mx_multifab* src,
                      int srccomp, int destcomp, int ncomp, const mx_bcrec* bcs,
                      int variable_type) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    if (lev != 0) return MX_ERR_UNSUPPORTED;  // above level 0: mx_adv_fill_patch_two_levels
    const bool is_vel = variable_type == 1;
    int ng[3];
    if ((rc = single_check(c, lev, mf, src, srccomp, destcomp, ncomp, bcs, is_vel, ng))) return rc;
    return single_run(c, lev, mf, src, srccomp, destcomp, ncomp, bcs, is_vel, ng);
}

// Maestro::FillPatch on level lev > 0 (MaestroFillData.cpp:48-66):
// amrex::FillPatchTwoLevels with cell_cons_interp
int mx_adv_fill_patch_two_levels(mx_adv_ctx* c, int lev, mx_multifab* mf,
                                 const mx_multifab* crse, const mx_multifab* fine, int srccomp,
                                 int destcomp, int ncomp, const mx_bcrec* bcs,
                                 int variable_type) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    if (lev < 1 || c->grids[lev - 1].empty()) return MX_ERR_ARG;
    if (ncomp > 32) return MX_ERR_ARG;  // the one-sided flags are 32-bit masks
    const bool is_vel = variable_type == 1;
    const int dim = c->p.dim;
    int ng[3];
    if ((rc = single_check(c, lev, mf, fine, srccomp, destcomp, ncomp, bcs, is_vel, ng)))
        return rc;
    const auto& cg = c->grids[lev - 1];
    if (!crse || !crse->fab || crse->nfab != (int)cg.size()) return MX_ERR_ARG;
    for (size_t b = 0; b < cg.size(); ++b)
        if (!covers(crse->fab[b], hbox(cg[b]), srccomp + ncomp)) return MX_ERR_BOX;
    int r[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) {
        const int nc = c->p.domain_hi[lev - 1][d] - c->p.domain_lo[lev - 1][d] + 1;
        const int nf = c->p.domain_hi[lev][d] - c->p.domain_lo[lev][d] + 1;
        if (nc <= 0 || nf % nc != 0) return MX_ERR_ARG;
        r[d] = nf / nc;
    }
    const HBox cdom = level_domain(c, lev - 1);
    // 1. fine patches: each FAB's box in the periodically grown fine domain
    //    that no fine valid box (or image) holds
    const auto fimgs = level_images(c, lev);
    const HBox fdg = periodic_grown(c, level_domain(c, lev));
    std::vector<std::pair<int, HBox>> patches;
    for (int b = 0; b < mf->nfab; ++b) {
        HBox x;
        if (!intersect(fab_box(mf->fab[b]), fdg, x)) continue;
        for (const HBox& q : uncovered(x, fimgs)) patches.push_back({b, q});
    }
    // 2. coarse patches, their coverage by the coarse level, the checks
    const auto cimgs = level_images(c, lev - 1);
    const HBox cdg = periodic_grown(c, cdom);
    std::vector<HBox> cps;
    for (const auto& pt : patches) {
        HBox cp;
        for (int d = 0; d < 3; ++d) {
            cp.lo[d] = floordiv_h(pt.second.lo[d], r[d]) - (d < dim ? 1 : 0);
            cp.hi[d] = floordiv_h(pt.second.hi[d], r[d]) + (d < dim ? 1 : 0);
        }
        HBox in;
        if (intersect(cp, cdg, in) && !uncovered(in, cimgs).empty()) return MX_ERR_BOX;
        cps.push_back(cp);
    }
    // the coarse patch holds the components at 0..ncomp-1, and that is the
    // component index its boundary functions see, as in the reference
    if ((rc = phys_check(c, cdom, cps, 0, ncomp, bcs, is_vel))) return rc;
    if (patches.empty()) return single_run(c, lev, mf, fine, srccomp, destcomp, ncomp, bcs, is_vel, ng);
    // coarse patch buffers in the workspace
    size_t need = 0;
    for (const HBox& cp : cps) need += (size_t)npts(cp) * ncomp * sizeof(double) + 256;
    if ((rc = ws_reserve(c, need))) return rc;
    std::vector<DA> bufs;
    for (const HBox& cp : cps) {
        DA a;
        const long long n0 = cp.hi[0] - cp.lo[0] + 1, n1 = cp.hi[1] - cp.lo[1] + 1,
                        n2 = cp.hi[2] - cp.lo[2] + 1;
        a.p = ws_take(c, (size_t)(n0 * n1 * n2 * ncomp));
        a.js = n0;
        a.ks = n0 * n1;
        a.ns = n0 * n1 * n2;
        a.b0 = cp.lo[0];
        a.b1 = cp.lo[1];
        a.b2 = cp.lo[2];
        bufs.push_back(a);
    }
    // the copy from the coarse valid boxes and their images
    {
        std::vector<CopyTask> tasks;
        for (size_t q = 0; q < cps.size(); ++q)
            for (const Image& im : cimgs) {
                HBox x;
                if (!intersect(cps[q], im.box, x)) continue;
                CopyTask T;
                for (int d = 0; d < 3; ++d) {
                    T.lo[d] = x.lo[d];
                    T.hi[d] = x.hi[d];
                    T.sh[d] = im.sh[d];
                }
                T.blk0 = 0;
                T.dst = bufs[q];
                T.src = to_da(crse->fab[im.src]);
                tasks.push_back(T);
            }
        std::vector<int> blk;
        if ((rc = flatten(tasks, blk))) return rc;
        if (!tasks.empty()) {
            Table tab(c);
            const size_t ot = tab.add(tasks);
            const size_t ob = tab.add(blk);
            if ((rc = tab.upload())) return rc;
            CopyArgs A{srccomp, 0};
            MX_LAUNCH(c, "k_fill_copy", k_fill_copy, dim3((unsigned)blk.size(), 1, (unsigned)ncomp),
                      dim3(NTF), tab.dev<CopyTask>(ot), tab.dev<int>(ob), A);
        }
    }
    // the coarse physical boundaries
    {
        std::vector<std::pair<DA, HBox>> fabs;
        for (size_t q = 0; q < cps.size(); ++q) fabs.push_back({bufs[q], cps[q]});
        if ((rc = phys_fill(c, cdom, fabs, 0, ncomp, bcs, is_vel))) return rc;
    }
    // 3. the interpolation into mf's fine patches
    {
        std::vector<InterpTask> tasks;
        for (size_t q = 0; q < patches.size(); ++q) {
            InterpTask T;
            for (int d = 0; d < 3; ++d) {
                T.lo[d] = patches[q].second.lo[d];
                T.hi[d] = patches[q].second.hi[d];
                T.clo[d] = cps[q].lo[d];
                T.chi[d] = cps[q].hi[d];
                T.one_lo[d] = T.one_hi[d] = 0u;
            }
            // setBC: the domain's BCRec on a side the coarse patch reaches
            for (int n = 0; n < ncomp; ++n)
                for (int d = 0; d < dim; ++d) {
                    const int blo = cps[q].lo[d] <= cdom.lo[d] ? bcs[n].lo[d] : MX_BC_INT_DIR;
                    const int bhi = cps[q].hi[d] >= cdom.hi[d] ? bcs[n].hi[d] : MX_BC_INT_DIR;
                    if (blo == MX_BC_EXT_DIR || blo == MX_BC_HOEXTRAP) T.one_lo[d] |= 1u << n;
                    if (bhi == MX_BC_EXT_DIR || bhi == MX_BC_HOEXTRAP) T.one_hi[d] |= 1u << n;
                }
            T.blk0 = 0;
            T.u = bufs[q];
            T.dst = to_da(mf->fab[patches[q].first]);
            tasks.push_back(T);
        }
        std::vector<int> blk;
        if ((rc = flatten(tasks, blk))) return rc;
        Table tab(c);
        const size_t ot = tab.add(tasks);
        const size_t ob = tab.add(blk);
        if ((rc = tab.upload())) return rc;
        InterpArgs A;
        for (int d = 0; d < 3; ++d) {
            A.r[d] = r[d];
            A.dlo[d] = cdom.lo[d];
            A.dhi[d] = cdom.hi[d];
        }
        A.dim = dim;
        A.ncomp = ncomp;
        A.dcomp = destcomp;
        MX_LAUNCH(c, "k_fill_interp", k_fill_interp, dim3((unsigned)blk.size()), dim3(NTF),
                  tab.dev<InterpTask>(ot), tab.dev<int>(ob), A);
    }
    // 4. the fine level's own valid data and physical boundaries
    return single_run(c, lev, mf, fine, srccomp, destcomp, ncomp, bcs, is_vel, ng);
}

}  // extern "C"
