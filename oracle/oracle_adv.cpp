// oracle_adv.cpp -- CPU oracle: MakeUtrans, VelPred, MakeEdgeScal, fluxes,
// Addw0, CelltoEdge, FillUmacGhost and the single-level FillBoundary.
//
// TEST INFRASTRUCTURE ONLY (see oracle_core.H).
//
// Each function restates the reference routine cited above it, loop by loop:
// the same temporaries, the same index boxes (obx = tile grown by one,
// growLo/growHi face boxes, nodal tile boxes), the same boundary overrides
// keyed on domain indices, and the same operand order in every expression.
// Where the reference repeats a body per direction, the body is written once
// with the direction as a parameter; direction- and dimension-specific
// differences of the reference are kept as explicit branches and commented.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>

#include "oracle_core.H"

namespace mxo {

int FluxRegOps(const Ctx& c, int lev, const std::vector<Arr> sflux[3], bool rhox,
               int start_comp, int num_comp);


void PPM(const Ctx& c, const IBox& bx, const Arr& s, const Arr& u,
         const Arr& v, const Arr& w, const Arr& Ip, const Arr& Im,
         const IBox& dom, const mx_bcrec* bcs, const double* dx, bool is_umac,
         int comp, int bccomp);
void Slope(int dir, int slope_order, const IBox& bx, const Arr& s,
           const Arr& sl, const IBox& dom, const mx_bcrec* bcs, int ncomp,
           int bc_start_comp);

namespace {

inline int e_(int d, int q) { return d == q ? 1 : 0; }

// Arr view starting at component `comp` (a one-component MultiFab::Copy)
inline Arr comp_view(const Arr& a, int comp) {
    Arr r = a;
    r.p = a.p + (long long)comp * a.ns;
    r.nc = a.nc - comp;
    return r;
}

inline bool contains(const IBox& b, int i, int j, int k) {
    return i >= b.lo[0] && i <= b.hi[0] && j >= b.lo[1] && j <= b.hi[1] &&
           k >= b.lo[2] && k <= b.hi[2];
}

}  // namespace

// ---------------------------------------------------------------------------
// Maestro::BCSetup (MaestroSetup.cpp:276-559): BCRec per component from
// phys_bc.
void BCSetup(const Ctx& c, std::vector<mx_bcrec>& bcs_s,
             std::vector<mx_bcrec>& bcs_u) {
    const mx_adv_params& p = c.p;
    bcs_s.assign(p.nscal, mx_bcrec{{0, 0, 0}, {0, 0, 0}});
    bcs_u.assign(3, mx_bcrec{{0, 0, 0}, {0, 0, 0}});
    for (int dir = 0; dir < p.dim; ++dir) {
        for (int side = 0; side < 2; ++side) {
            const int phys = side == 0 ? p.phys_bc_lo[dir] : p.phys_bc_hi[dir];
            int u_all = MX_BC_INT_DIR, u_dir = MX_BC_INT_DIR;
            int sc = MX_BC_INT_DIR, pi = MX_BC_INT_DIR;
            switch (phys) {
                case MX_PHYS_INTERIOR:
                    u_all = u_dir = sc = pi = MX_BC_INT_DIR;
                    break;
                case MX_PHYS_INFLOW:
                    u_all = u_dir = MX_BC_EXT_DIR;
                    sc = MX_BC_EXT_DIR;
                    pi = MX_BC_FOEXTRAP;
                    break;
                case MX_PHYS_OUTFLOW:
                    u_all = u_dir = MX_BC_FOEXTRAP;
                    sc = MX_BC_FOEXTRAP;
                    pi = MX_BC_EXT_DIR;
                    break;
                case MX_PHYS_SYMMETRY:
                    u_all = MX_BC_REFLECT_EVEN;
                    u_dir = MX_BC_REFLECT_ODD;
                    sc = MX_BC_REFLECT_EVEN;
                    pi = MX_BC_REFLECT_EVEN;
                    break;
                case MX_PHYS_SLIPWALL:
                    u_all = MX_BC_HOEXTRAP;
                    u_dir = MX_BC_EXT_DIR;
                    sc = MX_BC_HOEXTRAP;
                    pi = MX_BC_FOEXTRAP;
                    break;
                case MX_PHYS_NOSLIPWALL:
                    u_all = u_dir = MX_BC_EXT_DIR;
                    sc = MX_BC_HOEXTRAP;
                    pi = MX_BC_FOEXTRAP;
                    break;
                default:
                    break;
            }
            for (int comp = 0; comp < p.dim; ++comp) {
                int v = (comp == dir) ? u_dir : u_all;
                if (side == 0) bcs_u[comp].lo[dir] = v; else bcs_u[comp].hi[dir] = v;
            }
            for (int comp = 0; comp < p.nscal; ++comp) {
                int v = (comp == p.comp_pi) ? pi : sc;
                if (side == 0) bcs_s[comp].lo[dir] = v; else bcs_s[comp].hi[dir] = v;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// FillBoundary of one face-centred MultiFab on a single level (AMReX
// MultiFab::FillBoundary(periodicity)): every ghost face of a FAB that lies in
// the valid face box of some box of the level (after a periodic shift) is
// copied from there; valid faces are not touched.  A ghost face that lies in
// several valid face boxes (a face shared by two boxes, or a face on a
// periodic boundary) takes the value of the first source in the order
// (source box ascending, shift (sx, sy, sz) ascending from -1 to 1); when the
// sources agree, as they do for consistent data, the order is immaterial.
// A ghost face that no box of the level covers (a coarse-fine face of a finer
// level, which the reference fills by FillPatch from the coarser level) takes
// the value of the box's nearest valid face.
void FillBoundaryFace(const Ctx& c, int lev, const std::vector<Arr>& f,
                      int ndir, bool extrapolate) {
    const int dim = c.dim();
    const std::vector<IBox>& grids = c.grids[lev];
    const IBox dom = c.domain(lev);
    int nlen[3];
    for (int d = 0; d < 3; ++d) nlen[d] = dom.hi[d] - dom.lo[d] + 1;
    std::vector<std::array<int, 3>> shifts;
    for (int sx = -1; sx <= 1; ++sx)
        for (int sy = -1; sy <= 1; ++sy)
            for (int sz = -1; sz <= 1; ++sz) {
                int s[3] = {sx, sy, sz};
                bool ok = true;
                for (int d = 0; d < 3; ++d) {
                    if (s[d] != 0 && (d >= dim || !c.p.is_periodic[d])) ok = false;
                }
                if (ok) shifts.push_back({sx * nlen[0], sy * nlen[1], sz * nlen[2]});
            }
    for (size_t b = 0; b < grids.size(); ++b) {
        const Arr& dst = f[b];
        IBox valid = nodal(grids[b], ndir);
        IBox fabbox;
        for (int d = 0; d < 3; ++d) {
            fabbox.lo[d] = dst.b[d];
            fabbox.hi[d] = dst.e[d] - 1;
        }
        const long long fx = fabbox.hi[0] - fabbox.lo[0] + 1, fy = fabbox.hi[1] - fabbox.lo[1] + 1,
                        fz = fabbox.hi[2] - fabbox.lo[2] + 1;
        std::vector<char> covered((size_t)(fx * fy * fz), 0);
        auto at = [&](int i, int j, int k) {
            return (size_t)((i - fabbox.lo[0]) + fx * ((j - fabbox.lo[1]) + fy * (k - fabbox.lo[2])));
        };
        // sources in reverse order, so that the first source is written last
        for (size_t s1 = grids.size(); s1 > 0; --s1) {
            const size_t s = s1 - 1;
            IBox sv = nodal(grids[s], ndir);
            for (auto it = shifts.rbegin(); it != shifts.rend(); ++it) {
                const auto& sh = *it;
                IBox in;
                bool empty = false;
                for (int d = 0; d < 3; ++d) {
                    in.lo[d] = std::max(fabbox.lo[d], sv.lo[d] + sh[d]);
                    in.hi[d] = std::min(fabbox.hi[d], sv.hi[d] + sh[d]);
                    if (in.lo[d] > in.hi[d]) empty = true;
                }
                if (empty) continue;
                const Arr& src = f[s];
                for (int n = 0; n < dst.nc; ++n)
                    for (int k = in.lo[2]; k <= in.hi[2]; ++k)
                        for (int j = in.lo[1]; j <= in.hi[1]; ++j)
                            for (int i = in.lo[0]; i <= in.hi[0]; ++i) {
                                if (contains(valid, i, j, k)) continue;
                                dst(i, j, k, n) = src(i - sh[0], j - sh[1], k - sh[2], n);
                                covered[at(i, j, k)] = 1;
                            }
            }
        }
        // ghost faces no box of the level covers and that are not outside a
        // non-periodic domain side (coarse-fine faces of a finer level):
        // first-order extrapolation from the box's own nearest valid face
        for (int k = fabbox.lo[2]; k <= fabbox.hi[2]; ++k)
            for (int j = fabbox.lo[1]; j <= fabbox.hi[1]; ++j)
                for (int i = fabbox.lo[0]; i <= fabbox.hi[0]; ++i) {
                    if (!extrapolate) continue;
                    if (contains(valid, i, j, k) || covered[at(i, j, k)]) continue;
                    const int q[3] = {i, j, k};
                    bool phys = false;
                    for (int d = 0; d < dim; ++d) {
                        if (c.p.is_periodic[d]) continue;
                        const int hi = dom.hi[d] + (d == ndir ? 1 : 0);
                        if (q[d] < dom.lo[d] || q[d] > hi) phys = true;
                    }
                    if (phys) continue;
                    int r[3];
                    for (int d = 0; d < 3; ++d) r[d] = std::min(std::max(q[d], valid.lo[d]), valid.hi[d]);
                    for (int n = 0; n < dst.nc; ++n) dst(i, j, k, n) = dst(r[0], r[1], r[2], n);
                }
    }
}

// Maestro::FillUmacGhost (MaestroFillData.cpp:176-453), one level.
void FillUmacGhost(const Ctx& c, int lev, const std::vector<Arr> uedge[3]) {
    const int dim = c.dim();
    const IBox dom = c.domain(lev);
    const int* domlo = dom.lo;
    const int* domhi = dom.hi;
    const std::vector<IBox>& grids = c.grids[lev];

    for (size_t b = 0; b < grids.size(); ++b) {
        const Arr& umac = uedge[0][b];
        const Arr& vmac = uedge[1][b];
        Arr wmac;
        if (dim == 3) wmac = uedge[2][b];

        // x faces
        {
            IBox xbx = grow(nodal(grids[b], 0), dim, 1);
            const int bclo = c.p.phys_bc_lo[0], bchi = c.p.phys_bc_hi[0];
            for (int k = xbx.lo[2]; k <= xbx.hi[2]; ++k)
                for (int j = xbx.lo[1]; j <= xbx.hi[1]; ++j)
                    for (int i = xbx.lo[0]; i <= xbx.hi[0]; ++i) {
                        if (i == domlo[0] - 1) {
                            switch (bclo) {
                                case MX_PHYS_INFLOW:
                                    umac(i, j, k) = umac(i + 1, j, k);
                                    vmac(i, j, k) = 0.0;
                                    if (dim == 3) wmac(i, j, k) = 0.0;
                                    break;
                                case MX_PHYS_SLIPWALL:
                                case MX_PHYS_NOSLIPWALL:
                                    umac(i, j, k) = 0.0;
                                    vmac(i, j, k) = 0.0;
                                    if (dim == 3) wmac(i, j, k) = 0.0;
                                    break;
                                case MX_PHYS_OUTFLOW:
                                    umac(i, j, k) = umac(i + 1, j, k);
                                    vmac(i, j, k) = vmac(i + 1, j, k);
                                    if (dim == 3) wmac(i, j, k) = wmac(i + 1, j, k);
                                    break;
                                case MX_PHYS_SYMMETRY:
                                    umac(i, j, k) = -umac(i + 2, j, k);
                                    vmac(i, j, k) = vmac(i + 1, j, k);
                                    if (dim == 3) wmac(i, j, k) = wmac(i + 1, j, k);
                                    break;
                                default:
                                    break;
                            }
                        }
                        if (i == domhi[0] + 2) {
                            switch (bchi) {
                                case MX_PHYS_INFLOW:
                                    umac(i, j, k) = umac(i - 1, j, k);
                                    vmac(i - 1, j, k) = 0.0;
                                    if (dim == 3) wmac(i - 1, j, k) = 0.0;
                                    break;
                                case MX_PHYS_SLIPWALL:
                                case MX_PHYS_NOSLIPWALL:
                                    umac(i, j, k) = 0.0;
                                    vmac(i - 1, j, k) = 0.0;
                                    if (dim == 3) wmac(i - 1, j, k) = 0.0;
                                    break;
                                case MX_PHYS_OUTFLOW:
                                    umac(i, j, k) = umac(i - 1, j, k);
                                    vmac(i - 1, j, k) = vmac(i - 2, j, k);
                                    if (dim == 3) wmac(i - 1, j, k) = wmac(i - 2, j, k);
                                    break;
                                case MX_PHYS_SYMMETRY:
                                    umac(i, j, k) = -umac(i - 2, j, k);
                                    vmac(i - 1, j, k) = vmac(i - 2, j, k);
                                    if (dim == 3) wmac(i - 1, j, k) = wmac(i - 2, j, k);
                                    break;
                                default:
                                    break;
                            }
                        }
                    }
        }
        // y faces
        {
            IBox ybx = grow(nodal(grids[b], 1), dim, 1);
            const int bclo = c.p.phys_bc_lo[1], bchi = c.p.phys_bc_hi[1];
            for (int k = ybx.lo[2]; k <= ybx.hi[2]; ++k)
                for (int j = ybx.lo[1]; j <= ybx.hi[1]; ++j)
                    for (int i = ybx.lo[0]; i <= ybx.hi[0]; ++i) {
                        if (j == domlo[1] - 1) {
                            switch (bclo) {
                                case MX_PHYS_INFLOW:
                                    umac(i, j, k) = 0.0;
                                    vmac(i, j, k) = vmac(i, j + 1, k);
                                    if (dim == 3) wmac(i, j, k) = 0.0;
                                    break;
                                case MX_PHYS_SLIPWALL:
                                case MX_PHYS_NOSLIPWALL:
                                    umac(i, j, k) = 0.0;
                                    vmac(i, j, k) = 0.0;
                                    if (dim == 3) wmac(i, j, k) = 0.0;
                                    break;
                                case MX_PHYS_OUTFLOW:
                                    umac(i, j, k) = umac(i, j + 1, k);
                                    vmac(i, j, k) = vmac(i, j + 1, k);
                                    if (dim == 3) wmac(i, j, k) = wmac(i, j + 1, k);
                                    break;
                                case MX_PHYS_SYMMETRY:
                                    umac(i, j, k) = umac(i, j + 1, k);
                                    vmac(i, j, k) = -vmac(i, j + 2, k);
                                    if (dim == 3) wmac(i, j, k) = wmac(i, j + 1, k);
                                    break;
                                default:
                                    break;
                            }
                        }
                        if (j == domhi[1] + 2) {
                            switch (bchi) {
                                case MX_PHYS_INFLOW:
                                    umac(i, j - 1, k) = 0.0;
                                    vmac(i, j, k) = vmac(i, j - 1, k);
                                    if (dim == 3) wmac(i, j - 1, k) = 0.0;
                                    break;
                                case MX_PHYS_SLIPWALL:
                                case MX_PHYS_NOSLIPWALL:
                                    umac(i, j - 1, k) = 0.0;
                                    vmac(i, j, k) = 0.0;
                                    if (dim == 3) wmac(i, j - 1, k) = 0.0;
                                    break;
                                case MX_PHYS_OUTFLOW:
                                    umac(i, j - 1, k) = umac(i, j - 2, k);
                                    vmac(i, j, k) = vmac(i, j - 1, k);
                                    if (dim == 3) wmac(i, j - 1, k) = wmac(i, j - 2, k);
                                    break;
                                case MX_PHYS_SYMMETRY:
                                    umac(i, j - 1, k) = umac(i, j - 2, k);
                                    vmac(i, j, k) = -vmac(i, j - 2, k);
                                    if (dim == 3) wmac(i, j - 1, k) = wmac(i, j - 2, k);
                                    break;
                                default:
                                    break;
                            }
                        }
                    }
        }
        // z faces
        if (dim == 3) {
            IBox zbx = grow(nodal(grids[b], 2), dim, 1);
            const int bclo = c.p.phys_bc_lo[2], bchi = c.p.phys_bc_hi[2];
            for (int k = zbx.lo[2]; k <= zbx.hi[2]; ++k)
                for (int j = zbx.lo[1]; j <= zbx.hi[1]; ++j)
                    for (int i = zbx.lo[0]; i <= zbx.hi[0]; ++i) {
                        if (k == domlo[2] - 1) {
                            switch (bclo) {
                                case MX_PHYS_INFLOW:
                                    umac(i, j, k) = 0.0;
                                    vmac(i, j, k) = 0.0;
                                    wmac(i, j, k) = wmac(i, j, k + 1);
                                    break;
                                case MX_PHYS_SLIPWALL:
                                case MX_PHYS_NOSLIPWALL:
                                    umac(i, j, k) = 0.0;
                                    vmac(i, j, k) = 0.0;
                                    wmac(i, j, k) = 0.0;
                                    break;
                                case MX_PHYS_OUTFLOW:
                                    umac(i, j, k) = umac(i, j, k + 1);
                                    vmac(i, j, k) = vmac(i, j, k + 1);
                                    wmac(i, j, k) = wmac(i, j, k + 1);
                                    break;
                                case MX_PHYS_SYMMETRY:
                                    umac(i, j, k) = umac(i, j, k + 1);
                                    vmac(i, j, k) = vmac(i, j, k + 1);
                                    wmac(i, j, k) = -wmac(i, j, k + 2);
                                    break;
                                default:
                                    break;
                            }
                        }
                        if (k == domhi[2] + 2) {
                            switch (bchi) {
                                case MX_PHYS_INFLOW:
                                    umac(i, j, k - 1) = 0.0;
                                    vmac(i, j, k - 1) = 0.0;
                                    wmac(i, j, k) = wmac(i, j, k - 1);
                                    break;
                                case MX_PHYS_SLIPWALL:
                                case MX_PHYS_NOSLIPWALL:
                                    umac(i, j, k - 1) = 0.0;
                                    vmac(i, j, k - 1) = 0.0;
                                    wmac(i, j, k) = 0.0;
                                    break;
                                case MX_PHYS_OUTFLOW:
                                    umac(i, j, k - 1) = umac(i, j, k - 2);
                                    vmac(i, j, k - 1) = vmac(i, j, k - 2);
                                    wmac(i, j, k) = wmac(i, j, k - 1);
                                    break;
                                case MX_PHYS_SYMMETRY:
                                    umac(i, j, k - 1) = umac(i, j, k - 2);
                                    vmac(i, j, k - 1) = vmac(i, j, k - 2);
                                    wmac(i, j, k) = -wmac(i, j, k - 2);
                                    break;
                                default:
                                    break;
                            }
                        }
                    }
        }
    }
}

// FillBoundary of each face MultiFab then FillUmacGhost, as MakeUtrans and
// Addw0 do when finest_level == 0 (MaestroMakeUtrans.cpp:631-640), and as
// FillPatchUedge does on every level after its coarse-fine fill
// (MaestroFillData.cpp:824-831; there without the extrapolation of faces no
// box covers, which that fill has set).
void FillFaceGhosts(const Ctx& c, int lev, const std::vector<Arr> uedge[3], bool extrapolate) {
    for (int d = 0; d < c.dim(); ++d) FillBoundaryFace(c, lev, uedge[d], d, extrapolate);
    FillUmacGhost(c, lev, uedge);
}

// ---------------------------------------------------------------------------
// Maestro::MakeUtrans (MaestroMakeUtrans.cpp:6-648), one level.
int MakeUtrans(const Ctx& c, int lev, const std::vector<Arr>& utilde,
               const std::vector<Arr>& ufull, const std::vector<Arr> utrans[3],
               const std::vector<Arr> w0mac[3], const std::vector<Arr>& w0_cart,
               const std::vector<mx_bcrec>& bcs_u, bool fill_ghosts) {
    const int dim = c.dim();
    const IBox dom = c.domain(lev);
    const double* dx = c.p.dx[lev];
    const double dt2 = 0.5 * c.p.dt;
    const int* domlo = dom.lo;
    const int* domhi = dom.hi;
    const int ppm_type_local = c.p.ppm_type;
    const double rel_eps_local = c.p.rel_eps;
    const std::vector<IBox>& grids = c.grids[lev];

    TmpMF Ip, Im;
    Ip.define(grids, dim, -1, dim, 2);
    Im.define(grids, dim, -1, dim, 2);

    for (int d = 0; d < dim; ++d) {
        const int bclo = c.p.phys_bc_lo[d];
        const int bchi = c.p.phys_bc_hi[d];
        const int e0 = e_(d, 0), e1 = e_(d, 1), e2 = e_(d, 2);
        for (size_t b = 0; b < grids.size(); ++b) {
            const IBox obx = grow(grids[b], dim, 1);
            const IBox nbx = nodal(grids[b], d);
            const Arr& ut = utilde[b];
            const Arr& uf = ufull[b];
            const Arr& Ipa = Ip[b];
            const Arr& Ima = Im[b];

            if (ppm_type_local == 0) {
                // u_mf/v_mf/w_mf are copies of utilde's components
                Slope(d, c.p.slope_order, obx, comp_view(ut, d), Ipa, dom,
                      bcs_u.data(), 1, d);
            } else {
                // u_mf/v_mf/w_mf are copies of ufull's components
                Arr uu = comp_view(uf, 0), vv = comp_view(uf, 1), ww;
                if (dim == 3) ww = comp_view(uf, 2);
                PPM(c, obx, ut, uu, vv, ww, Ipa, Ima, dom, bcs_u.data(), dx,
                    false, d, d);
            }

            const Arr& tr = utrans[d][b];
            Arr w0m;
            if (dim == 3 && c.p.spherical) w0m = w0mac[d][b];
            Arr w0c;
            if (d == dim - 1 && !c.p.spherical) w0c = w0_cart[b];

            par_for(nbx, [&](int i, int j, int k) {
                double ul = 0.0, ur = 0.0;
                if (ppm_type_local == 0) {
                    ul = ut(i - e0, j - e1, k - e2, d) +
                         (0.5 - (dt2 / dx[d]) * mx(0.0, uf(i - e0, j - e1, k - e2, d))) *
                             Ipa(i - e0, j - e1, k - e2, 0);
                    ur = ut(i, j, k, d) -
                         (0.5 + (dt2 / dx[d]) * mn(0.0, uf(i, j, k, d))) * Ipa(i, j, k, 0);
                } else {
                    ul = Ipa(i - e0, j - e1, k - e2, d);
                    ur = Ima(i, j, k, d);
                }
                const int pos = (d == 0) ? i : (d == 1) ? j : k;
                if (pos == domlo[d]) {
                    switch (bclo) {
                        case MX_PHYS_INFLOW:
                            ul = ut(i - e0, j - e1, k - e2, d);
                            ur = ut(i - e0, j - e1, k - e2, d);
                            break;
                        case MX_PHYS_SLIPWALL:
                        case MX_PHYS_NOSLIPWALL:
                        case MX_PHYS_SYMMETRY:
                            if (dim == 2 && d == 1) {
                                // 2-D y: "vry = 0.0; vry = 0.0;" (vly kept)
                                ur = 0.0;
                                ur = 0.0;
                            } else {
                                ul = 0.0;
                                ur = 0.0;
                            }
                            break;
                        case MX_PHYS_OUTFLOW:
                            ul = mn(ur, 0.0);
                            ur = ul;
                            break;
                        default:
                            break;
                    }
                } else if (pos == domhi[d] + 1) {
                    switch (bchi) {
                        case MX_PHYS_INFLOW:
                            if (dim == 2) {
                                // 2-D takes the last valid cell
                                ul = ut(i, j, k, d);
                                ur = ut(i, j, k, d);
                            } else {
                                ul = ut(i + e0, j + e1, k + e2, d);
                                ur = ut(i + e0, j + e1, k + e2, d);
                            }
                            break;
                        case MX_PHYS_SLIPWALL:
                        case MX_PHYS_NOSLIPWALL:
                        case MX_PHYS_SYMMETRY:
                            ul = 0.0;
                            ur = 0.0;
                            break;
                        case MX_PHYS_OUTFLOW:
                            ul = mx(ul, 0.0);
                            ur = ul;
                            break;
                        default:
                            break;
                    }
                }

                double res;
                bool use_w = false;
                double w = 0.0;
                if (dim == 3 && c.p.spherical) {
                    use_w = true;
                    w = w0m(i, j, k);
                } else if (d == dim - 1) {
                    use_w = true;
                    w = w0c(i, j, k, dim - 1);
                }
                if (use_w) {
                    bool test = (ul + w <= 0.0 && ur + w >= 0.0) ||
                                (ab(ul + ur + 2.0 * w) < rel_eps_local);
                    res = 0.5 * (ul + ur) + w > 0.0 ? ul : ur;
                    res = test ? 0.0 : res;
                } else {
                    bool test = (ul <= 0.0 && ur >= 0.0) ||
                                (ab(ul + ur) < rel_eps_local);
                    res = 0.5 * (ul + ur) > 0.0 ? ul : ur;
                    res = test ? 0.0 : res;
                }
                tr(i, j, k) = res;
            });
        }
    }

    // finest_level == 0; with finer levels the caller runs AverageDownFaces
    // and FillPatchUedge over the hierarchy instead (:641-647)
    if (fill_ghosts) FillFaceGhosts(c, lev, utrans, true);
    return MX_OK;
}

// ---------------------------------------------------------------------------
// Maestro::Addw0 (MaestroFill3dData.cpp:357-461), one level.
int Addw0(const Ctx& c, int lev, const std::vector<Arr> uedge[3],
          const std::vector<Arr> w0mac[3], const BaseArr& w0, double mult, bool fill_ghosts) {
    const int dim = c.dim();
    const std::vector<IBox>& grids = c.grids[lev];
    for (size_t b = 0; b < grids.size(); ++b) {
        if (!c.p.spherical) {
            const int d = dim - 1;
            IBox bx = nodal(grids[b], d);
            for (int t = 0; t < dim; ++t)
                if (t != d) bx = grow_dir(bx, t, 1);
            const Arr& ve = uedge[d][b];
            par_for(bx, [&](int i, int j, int k) {
                const int r = (dim == 2) ? j : k;
                ve(i, j, k) += mult * w0(lev, r);
            });
        } else {
            if (dim != 3) return MX_ERR_UNSUPPORTED;
            for (int d = 0; d < 3; ++d) {
                IBox bx = nodal(grids[b], d);
                const Arr& ue = uedge[d][b];
                const Arr& wm = w0mac[d][b];
                par_for(bx, [&](int i, int j, int k) { ue(i, j, k) += mult * wm(i, j, k); });
            }
        }
    }
    // finest_level == 0 (MaestroFill3dData.cpp:424-448); with finer levels
    // the caller runs AverageDownFaces and FillPatchUedge (:449-455)
    if (fill_ghosts) FillFaceGhosts(c, lev, uedge, true);
    return MX_OK;
}

// ---------------------------------------------------------------------------
// Maestro::VelPred (MaestroVelPred.cpp:5-292) with VelPredInterface,
// VelPredTransverse and VelPredVelocities (2-D :296-731, 3-D :735-1915).
int VelPred(const Ctx& c, int lev, const std::vector<Arr>& utilde,
            const std::vector<Arr>& ufull, const std::vector<Arr> utrans[3],
            const std::vector<Arr> umac[3], const std::vector<Arr> w0mac[3],
            const std::vector<Arr>& force, const std::vector<Arr>& w0_cart,
            const std::vector<mx_bcrec>& bcs_u) {
    const int dim = c.dim();
    const IBox dom = c.domain(lev);
    const double* dx = c.p.dx[lev];
    const int* domlo = dom.lo;
    const int* domhi = dom.hi;
    const double dt = c.p.dt;
    const double dt2 = 0.5 * dt;
    const double dt4 = 0.25 * dt;
    const double dt6 = dt / 6.0;
    const double rel_eps_local = c.p.rel_eps;
    const int ppm_type = c.p.ppm_type;
    const int ppm_trace_forces = c.p.ppm_trace_forces;
    const std::vector<IBox>& grids = c.grids[lev];

    if (dim == 2 && ppm_trace_forces == 1) {
        // the 2-D reference writes the y force trace into Ipv/Imv and then
        // reads the never-written Ipfy/Imfy (MaestroVelPred.cpp:128-133,556)
        return MX_ERR_UNSUPPORTED;
    }
    for (int d = 0; d < dim; ++d) {
        for (int side = 0; side < 2; ++side) {
            int pb = side == 0 ? c.p.phys_bc_lo[d] : c.p.phys_bc_hi[d];
            if (pb < MX_PHYS_INTERIOR || pb > MX_PHYS_NOSLIPWALL) return MX_ERR_BC;
        }
    }

    // Ip*/Im* of each velocity component; Ipf*/Imf* of each force component
    TmpMF Ipc[3], Imc[3], Ipf[3], Imf[3];
    TmpMF ul[3], ur[3], uimh[3];
    TmpMF T[3][3];  // T[d][t]: transverse state on d-faces corrected by t
    for (int n = 0; n < dim; ++n) {
        Ipc[n].define(grids, dim, -1, dim, 1);
        Imc[n].define(grids, dim, -1, dim, 1);
        Ipf[n].define(grids, dim, -1, dim, 1);
        Imf[n].define(grids, dim, -1, dim, 1);
        ul[n].define(grids, dim, -1, dim, 1);
        ur[n].define(grids, dim, -1, dim, 1);
        uimh[n].define(grids, dim, -1, dim, 1);
    }
    if (dim == 3) {
        for (int d = 0; d < 3; ++d)
            for (int t = 0; t < 3; ++t)
                if (t != d) T[d][t].define(grids, dim, -1, 1, 1);
    }

    for (size_t b = 0; b < grids.size(); ++b) {
        const IBox obx = grow(grids[b], dim, 1);
        const Arr& ut = utilde[b];
        const Arr& uf = ufull[b];
        const Arr& fo = force[b];
        Arr uu, vv, ww;
        if (ppm_type != 0) {
            uu = comp_view(uf, 0);
            vv = comp_view(uf, 1);
            if (dim == 3) ww = comp_view(uf, 2);
        }

        // slopes go to Ipu (x), Imv (y), Imw (z) when ppm_type == 0
        for (int n = 0; n < dim; ++n) {
            if (ppm_type == 0) {
                const Arr& dst = (n == 0) ? Ipc[0][b] : Imc[n][b];
                Slope(n, c.p.slope_order, obx, ut, dst, dom, bcs_u.data(), dim, 0);
            } else {
                PPM(c, obx, ut, uu, vv, ww, Ipc[n][b], Imc[n][b], dom,
                    bcs_u.data(), dx, false, n, n);
                if (ppm_trace_forces == 1) {
                    PPM(c, obx, fo, uu, vv, ww, Ipf[n][b], Imf[n][b], dom,
                        bcs_u.data(), dx, false, n, n);
                }
            }
        }

        // ---- VelPredInterface
        int bad_bc = 0;
        for (int d = 0; d < dim; ++d) {
            const int e0 = e_(d, 0), e1 = e_(d, 1), e2 = e_(d, 2);
            const IBox mbx = grow_lo(obx, d, -1);
            const int bclo = c.p.phys_bc_lo[d], bchi = c.p.phys_bc_hi[d];
            const Arr& L = ul[d][b];
            const Arr& R = ur[d][b];
            const Arr& U = uimh[d][b];
            const Arr& tr = utrans[d][b];
            const Arr& SL = (ppm_type == 0) ? ((d == 0) ? Ipc[0][b] : Imc[d][b]) : Ipc[0][b];
            const double hd = dx[d];
            par_for(mbx, [&](int i, int j, int k) {
                const int im = i - e0, jm = j - e1, km = k - e2;
                if (ppm_type == 0) {
                    if (dim == 3) {
                        double maxu = 0.5 - dt2 * mx(0.0, uf(im, jm, km, d)) / hd;
                        double minu = 0.5 + dt2 * mn(0.0, uf(i, j, k, d)) / hd;
                        for (int n = 0; n < 3; ++n) {
                            L(i, j, k, n) = ut(im, jm, km, n) + maxu * SL(im, jm, km, n);
                            R(i, j, k, n) = ut(i, j, k, n) - minu * SL(i, j, k, n);
                        }
                    } else {
                        double maxu = mx(0.0, uf(im, jm, km, d));
                        double minu = mn(0.0, uf(i, j, k, d));
                        for (int n = 0; n < 2; ++n) {
                            L(i, j, k, n) = ut(im, jm, km, n) +
                                            (0.5 - (dt2 / hd) * maxu) * SL(im, jm, km, n);
                            R(i, j, k, n) = ut(i, j, k, n) -
                                            (0.5 + (dt2 / hd) * minu) * SL(i, j, k, n);
                        }
                    }
                } else {
                    for (int n = 0; n < dim; ++n) {
                        L(i, j, k, n) = Ipc[n][b](im, jm, km, d);
                        R(i, j, k, n) = Imc[n][b](i, j, k, d);
                    }
                }

                const int pos = (d == 0) ? i : (d == 1) ? j : k;
                if (pos == domlo[d]) {
                    switch (bclo) {
                        case MX_PHYS_INFLOW:
                            for (int n = 0; n < dim; ++n) {
                                L(i, j, k, n) = ut(im, jm, km, n);
                                R(i, j, k, n) = ut(im, jm, km, n);
                            }
                            break;
                        case MX_PHYS_SLIPWALL:
                        case MX_PHYS_SYMMETRY:
                            L(i, j, k, d) = 0.0;
                            R(i, j, k, d) = 0.0;
                            for (int n = 0; n < dim; ++n)
                                if (n != d) L(i, j, k, n) = R(i, j, k, n);
                            break;
                        case MX_PHYS_NOSLIPWALL:
                            for (int n = 0; n < dim; ++n) {
                                L(i, j, k, n) = 0.0;
                                R(i, j, k, n) = 0.0;
                            }
                            break;
                        case MX_PHYS_OUTFLOW:
                            R(i, j, k, d) = mn(R(i, j, k, d), 0.0);
                            if (dim == 2 && d == 0) {
                                // 2-D x: "urx(n) = ulx(n)" (MaestroVelPred.cpp:376-381)
                                for (int n = 0; n < dim; ++n) R(i, j, k, n) = L(i, j, k, n);
                            } else {
                                for (int n = 0; n < dim; ++n) L(i, j, k, n) = R(i, j, k, n);
                            }
                            break;
                        case MX_PHYS_INTERIOR:
                            break;
                        default:
                            bad_bc = 1;
                    }
                } else if (pos == domhi[d] + 1) {
                    switch (bchi) {
                        case MX_PHYS_INFLOW:
                            for (int n = 0; n < dim; ++n) {
                                L(i, j, k, n) = ut(i, j, k, n);
                                R(i, j, k, n) = ut(i, j, k, n);
                            }
                            break;
                        case MX_PHYS_SLIPWALL:
                        case MX_PHYS_SYMMETRY:
                            L(i, j, k, d) = 0.0;
                            R(i, j, k, d) = 0.0;
                            for (int n = 0; n < dim; ++n)
                                if (n != d) R(i, j, k, n) = L(i, j, k, n);
                            break;
                        case MX_PHYS_NOSLIPWALL:
                            for (int n = 0; n < dim; ++n) {
                                L(i, j, k, n) = 0.0;
                                R(i, j, k, n) = 0.0;
                            }
                            break;
                        case MX_PHYS_OUTFLOW:
                            L(i, j, k, d) = mx(L(i, j, k, d), 0.0);
                            for (int n = 0; n < dim; ++n) R(i, j, k, n) = L(i, j, k, n);
                            break;
                        case MX_PHYS_INTERIOR:
                            break;
                        default:
                            bad_bc = 1;
                    }
                }

                for (int n = 0; n < dim; ++n) {
                    if (n == d) continue;
                    double t = tr(i, j, k);
                    U(i, j, k, n) = t > 0.0 ? L(i, j, k, n) : R(i, j, k, n);
                    U(i, j, k, n) = ab(t) < rel_eps_local
                                        ? 0.5 * (L(i, j, k, n) + R(i, j, k, n))
                                        : U(i, j, k, n);
                }
            });
            (void)SL;
        }
        if (bad_bc) return MX_ERR_BC;

        // ---- VelPredTransverse (3-D)
        if (dim == 3) {
            for (int d = 0; d < 3; ++d) {
                for (int t = 0; t < 3; ++t) {
                    if (t == d) continue;
                    const int o = 3 - d - t;  // component of the state
                    IBox imhbox = grow_hi(grow_dir(grids[b], o, 1), d, 1);
                    const int ed0 = e_(d, 0), ed1 = e_(d, 1), ed2 = e_(d, 2);
                    const int et0 = e_(t, 0), et1 = e_(t, 1), et2 = e_(t, 2);
                    const int bclo = c.p.phys_bc_lo[d], bchi = c.p.phys_bc_hi[d];
                    const Arr& L = ul[d][b];
                    const Arr& R = ur[d][b];
                    const Arr& trt = utrans[t][b];
                    const Arr& trd = utrans[d][b];
                    const Arr& Ut = uimh[t][b];
                    const Arr& out = T[d][t][b];
                    const double ht = dx[t];
                    par_for(imhbox, [&](int i, int j, int k) {
                        const int im = i - ed0, jm = j - ed1, km = k - ed2;
                        double l = L(i, j, k, o) -
                                   (dt6 / ht) *
                                       (trt(im + et0, jm + et1, km + et2) + trt(im, jm, km)) *
                                       (Ut(im + et0, jm + et1, km + et2, o) - Ut(im, jm, km, o));
                        double r = R(i, j, k, o) -
                                   (dt6 / ht) * (trt(i + et0, j + et1, k + et2) + trt(i, j, k)) *
                                       (Ut(i + et0, j + et1, k + et2, o) - Ut(i, j, k, o));
                        const int pos = (d == 0) ? i : (d == 1) ? j : k;
                        if (pos == domlo[d]) {
                            switch (bclo) {
                                case MX_PHYS_INFLOW:
                                    l = ut(im, jm, km, o);
                                    r = ut(im, jm, km, o);
                                    break;
                                case MX_PHYS_SLIPWALL:
                                case MX_PHYS_SYMMETRY:
                                case MX_PHYS_OUTFLOW:
                                    l = r;
                                    break;
                                case MX_PHYS_NOSLIPWALL:
                                    l = 0.0;
                                    r = 0.0;
                                    break;
                                default:
                                    break;
                            }
                        } else if (pos == domhi[d] + 1) {
                            switch (bchi) {
                                case MX_PHYS_INFLOW:
                                    l = ut(i, j, k, o);
                                    r = ut(i, j, k, o);
                                    break;
                                case MX_PHYS_SLIPWALL:
                                case MX_PHYS_SYMMETRY:
                                case MX_PHYS_OUTFLOW:
                                    r = l;
                                    break;
                                case MX_PHYS_NOSLIPWALL:
                                    l = 0.0;
                                    r = 0.0;
                                    break;
                                default:
                                    break;
                            }
                        }
                        double tv = trd(i, j, k);
                        double res = tv > 0.0 ? l : r;
                        res = ab(tv) < rel_eps_local ? 0.5 * (l + r) : res;
                        out(i, j, k) = res;
                    });
                }
            }
        }

        // ---- VelPredVelocities
        for (int d = 0; d < dim; ++d) {
            const IBox nbx = nodal(grids[b], d);
            const int e0 = e_(d, 0), e1 = e_(d, 1), e2 = e_(d, 2);
            const int bclo = c.p.phys_bc_lo[d], bchi = c.p.phys_bc_hi[d];
            const Arr& L = ul[d][b];
            const Arr& R = ur[d][b];
            const Arr& out = umac[d][b];
            int tdir[2];
            int nt = 0;
            for (int t = 0; t < dim; ++t)
                if (t != d) tdir[nt++] = t;
            Arr w0m;
            if (dim == 3 && c.p.spherical) w0m = w0mac[d][b];
            Arr w0c;
            if (d == dim - 1 && !c.p.spherical) w0c = w0_cart[b];
            const Arr& IpfA = Ipf[d][b];
            const Arr& ImfA = Imf[d][b];

            par_for(nbx, [&](int i, int j, int k) {
                const int im = i - e0, jm = j - e1, km = k - e2;
                double fl = ppm_trace_forces == 0 ? fo(im, jm, km, d) : IpfA(im, jm, km, d);
                double fr = ppm_trace_forces == 0 ? fo(i, j, k, d) : ImfA(i, j, k, d);
                double vl = L(i, j, k, d);
                double vr = R(i, j, k, d);
                for (int q = 0; q < nt; ++q) {
                    const int t = tdir[q];
                    const int et0 = e_(t, 0), et1 = e_(t, 1), et2 = e_(t, 2);
                    const Arr& trt = utrans[t][b];
                    const double ht = dx[t];
                    if (dim == 3) {
                        const Arr& Tt = T[t][3 - d - t][b];
                        vl = vl - (dt4 / ht) *
                                      (trt(im + et0, jm + et1, km + et2) + trt(im, jm, km)) *
                                      (Tt(im + et0, jm + et1, km + et2) - Tt(im, jm, km));
                        vr = vr - (dt4 / ht) * (trt(i + et0, j + et1, k + et2) + trt(i, j, k)) *
                                      (Tt(i + et0, j + et1, k + et2) - Tt(i, j, k));
                    } else {
                        const Arr& Ut = uimh[t][b];
                        vl = vl - (dt4 / ht) *
                                      (trt(im + et0, jm + et1, km + et2) + trt(im, jm, km)) *
                                      (Ut(im + et0, jm + et1, km + et2, d) - Ut(im, jm, km, d));
                        vr = vr - (dt4 / ht) * (trt(i + et0, j + et1, k + et2) + trt(i, j, k)) *
                                      (Ut(i + et0, j + et1, k + et2, d) - Ut(i, j, k, d));
                    }
                }
                double umacl = vl + dt2 * fl;
                double umacr = vr + dt2 * fr;

                double res;
                bool use_w = false;
                double w = 0.0;
                if (dim == 3 && c.p.spherical) {
                    use_w = true;
                    w = w0m(i, j, k);
                } else if (d == dim - 1) {
                    use_w = true;
                    w = w0c(i, j, k, dim - 1);
                }
                if (use_w) {
                    bool test = (umacl + w <= 0.0 && umacr + w >= 0.0) ||
                                (ab(umacl + umacr + 2.0 * w) < rel_eps_local);
                    res = 0.5 * (umacl + umacr) + w > 0.0 ? umacl : umacr;
                    res = test ? 0.0 : res;
                } else {
                    bool test = (umacl <= 0.0 && umacr >= 0.0) ||
                                (ab(umacl + umacr) < rel_eps_local);
                    res = 0.5 * (umacl + umacr) > 0.0 ? umacl : umacr;
                    res = test ? 0.0 : res;
                }

                const int pos = (d == 0) ? i : (d == 1) ? j : k;
                if (pos == domlo[d]) {
                    switch (bclo) {
                        case MX_PHYS_INFLOW:
                            res = ut(im, jm, km, d);
                            break;
                        case MX_PHYS_SLIPWALL:
                        case MX_PHYS_SYMMETRY:
                        case MX_PHYS_NOSLIPWALL:
                            res = 0.0;
                            break;
                        case MX_PHYS_OUTFLOW:
                            res = mn(umacr, 0.0);
                            break;
                        default:
                            break;
                    }
                } else if (pos == domhi[d] + 1) {
                    switch (bchi) {
                        case MX_PHYS_INFLOW:
                            res = ut(i, j, k, d);
                            break;
                        case MX_PHYS_SLIPWALL:
                        case MX_PHYS_SYMMETRY:
                        case MX_PHYS_NOSLIPWALL:
                            res = 0.0;
                            break;
                        case MX_PHYS_OUTFLOW:
                            res = mx(umacl, 0.0);
                            break;
                        default:
                            break;
                    }
                }
                out(i, j, k) = res;
            });
        }
    }
    // AverageDownFaces(umac) is a no-op on a single level
    return MX_OK;
}

// ---------------------------------------------------------------------------
// Maestro::MakeEdgeScal (MaestroMakeEdgeScalars.cpp:6-302) with
// MakeEdgeScalPredictor / Transverse / Edges / MakeDivU.
int MakeEdgeScal(const Ctx& c, int lev, const std::vector<Arr>& state,
                 const std::vector<Arr> sedge[3], const std::vector<Arr> umac[3],
                 const std::vector<Arr>& force, bool is_vel,
                 const mx_bcrec* bcs, int start_scomp, int start_bccomp,
                 int num_comp, bool is_conservative) {
    const int dim = c.dim();
    const IBox dom = c.domain(lev);
    const double* dx = c.p.dx[lev];
    const int* domlo = dom.lo;
    const int* domhi = dom.hi;
    const double dt = c.p.dt;
    const double dt2 = 0.5 * dt;
    const double dt3 = dt / 3.0;
    const double dt4 = 0.25 * dt;
    const double dt6 = dt / 6.0;
    const double rel_eps_local = c.p.rel_eps;
    const int ppm_type = c.p.ppm_type;
    const int ppm_trace_forces = c.p.ppm_trace_forces;
    const std::vector<IBox>& grids = c.grids[lev];

    TmpMF Ip, Im, Ipf, Imf;
    Ip.define(grids, dim, -1, dim, 1);
    Im.define(grids, dim, -1, dim, 1);
    Ipf.define(grids, dim, -1, dim, 1);
    Imf.define(grids, dim, -1, dim, 1);
    TmpMF sl[3], sr[3], simh[3], slopez, divu;
    for (int d = 0; d < dim; ++d) {
        sl[d].define(grids, dim, -1, 1, 1, 0.0);
        sr[d].define(grids, dim, -1, 1, 1, 0.0);
        simh[d].define(grids, dim, -1, 1, 1, 0.0);
    }
    TmpMF T[3][3];
    if (dim == 3) {
        slopez.define(grids, dim, -1, 1, 1);
        divu.define(grids, dim, -1, 1, 1);
        for (int d = 0; d < 3; ++d)
            for (int t = 0; t < 3; ++t)
                if (t != d) T[d][t].define(grids, dim, -1, 1, 1, 0.0);
    }

    for (int scomp = start_scomp; scomp < start_scomp + num_comp; ++scomp) {
        const int comp = scomp;
        const int bccomp = start_bccomp + scomp - start_scomp;

        for (size_t b = 0; b < grids.size(); ++b) {
            const IBox obx = grow(grids[b], dim, 1);
            const Arr& s = state[b];
            Arr uu = umac[0][b], vv = umac[1][b], ww;
            if (dim == 3) ww = umac[2][b];

            if (dim == 3 && is_conservative) {
                // MakeDivU (MaestroMakeEdgeScalars.cpp:648-660)
                const Arr& dv = divu[b];
                par_for(obx, [&](int i, int j, int k) {
                    dv(i, j, k) = uu(i + 1, j, k) - uu(i, j, k) + vv(i, j + 1, k) -
                                  vv(i, j, k) + ww(i, j, k + 1) - ww(i, j, k);
                    dv(i, j, k) /= dx[0];
                });
            }

            if (ppm_type == 0) {
                Arr sc = comp_view(s, scomp);
                Slope(0, c.p.slope_order, obx, sc, Ip[b], dom, bcs, 1, bccomp);
                Slope(1, c.p.slope_order, obx, sc, Im[b], dom, bcs, 1, bccomp);
                if (dim == 3) Slope(2, c.p.slope_order, obx, sc, slopez[b], dom, bcs, 1, bccomp);
            } else {
                PPM(c, obx, s, uu, vv, ww, Ip[b], Im[b], dom, bcs, dx, true, scomp, bccomp);
                if (ppm_trace_forces == 1) {
                    PPM(c, obx, force[b], uu, vv, ww, Ipf[b], Imf[b], dom, bcs, dx,
                        true, scomp, bccomp);
                }
            }

            // ---- MakeEdgeScalPredictor
            for (int d = 0; d < dim; ++d) {
                const IBox mbx = grow_lo(obx, d, -1);
                const int e0 = e_(d, 0), e1 = e_(d, 1), e2 = e_(d, 2);
                const int bclo = bcs[bccomp].lo[d], bchi = bcs[bccomp].hi[d];
                const Arr& L = sl[d][b];
                const Arr& R = sr[d][b];
                const Arr& H = simh[d][b];
                const Arr& um = umac[d][b];
                const Arr& SLP = (d == 0) ? Ip[b] : (d == 1) ? Im[b] : slopez[b];
                const double h = dx[d];
                const double thr = (dim == 3) ? 0.0 : rel_eps_local;
                par_for(mbx, [&](int i, int j, int k) {
                    const int im = i - e0, jm = j - e1, km = k - e2;
                    if (ppm_type == 0) {
                        if (dim == 3) {
                            L(i, j, k) = s(im, jm, km, comp) +
                                         0.5 * (1.0 - dt * um(i, j, k) / h) * SLP(im, jm, km, 0);
                            R(i, j, k) = s(i, j, k, comp) -
                                         0.5 * (1.0 + dt * um(i, j, k) / h) * SLP(i, j, k, 0);
                        } else {
                            L(i, j, k) = s(im, jm, km, comp) +
                                         (0.5 - dt2 * um(i, j, k) / h) * SLP(im, jm, km, 0);
                            R(i, j, k) = s(i, j, k, comp) -
                                         (0.5 + dt2 * um(i, j, k) / h) * SLP(i, j, k, 0);
                        }
                    } else {
                        L(i, j, k) = Ip[b](im, jm, km, d);
                        R(i, j, k) = Im[b](i, j, k, d);
                    }
                    const int pos = (d == 0) ? i : (d == 1) ? j : k;
                    if (pos == domlo[d]) {
                        if (bclo == MX_BC_EXT_DIR) {
                            L(i, j, k) = s(im, jm, km, comp);
                            R(i, j, k) = s(im, jm, km, comp);
                        } else if (bclo == MX_BC_FOEXTRAP || bclo == MX_BC_HOEXTRAP) {
                            if (is_vel && comp == d) R(i, j, k) = mn(R(i, j, k), 0.0);
                            L(i, j, k) = R(i, j, k);
                        } else if (bclo == MX_BC_REFLECT_EVEN) {
                            L(i, j, k) = R(i, j, k);
                        } else if (bclo == MX_BC_REFLECT_ODD) {
                            L(i, j, k) = 0.0;
                            R(i, j, k) = 0.0;
                        }
                    } else if (pos == domhi[d] + 1) {
                        if (bchi == MX_BC_EXT_DIR) {
                            L(i, j, k) = s(i, j, k, comp);
                            R(i, j, k) = s(i, j, k, comp);
                        } else if (bchi == MX_BC_FOEXTRAP || bchi == MX_BC_HOEXTRAP) {
                            if (is_vel && comp == d) L(i, j, k) = mx(L(i, j, k), 0.0);
                            R(i, j, k) = L(i, j, k);
                        } else if (bchi == MX_BC_REFLECT_EVEN) {
                            R(i, j, k) = L(i, j, k);
                        } else if (bchi == MX_BC_REFLECT_ODD) {
                            L(i, j, k) = 0.0;
                            R(i, j, k) = 0.0;
                        }
                    }
                    H(i, j, k) = (um(i, j, k) > 0.0) ? L(i, j, k) : R(i, j, k);
                    H(i, j, k) = (ab(um(i, j, k)) > thr) ? H(i, j, k)
                                                         : 0.5 * (L(i, j, k) + R(i, j, k));
                });
            }

            // ---- MakeEdgeScalTransverse (3-D)
            if (dim == 3) {
                for (int d = 0; d < 3; ++d) {
                    for (int t = 0; t < 3; ++t) {
                        if (t == d) continue;
                        const int o = 3 - d - t;
                        IBox imhbox = grow_hi(grow_dir(grids[b], o, 1), d, 1);
                        const int ed0 = e_(d, 0), ed1 = e_(d, 1), ed2 = e_(d, 2);
                        const int et0 = e_(t, 0), et1 = e_(t, 1), et2 = e_(t, 2);
                        const int bclo = bcs[bccomp].lo[d], bchi = bcs[bccomp].hi[d];
                        const Arr& L = sl[d][b];
                        const Arr& R = sr[d][b];
                        const Arr& ut = umac[t][b];
                        const Arr& ud = umac[d][b];
                        const Arr& Ht = simh[t][b];
                        const Arr& dv = divu[b];
                        const Arr& out = T[d][t][b];
                        const double ht = dx[t];
                        par_for(imhbox, [&](int i, int j, int k) {
                            const int im = i - ed0, jm = j - ed1, km = k - ed2;
                            double l, r;
                            if (is_conservative) {
                                l = L(i, j, k) -
                                    (dt3 / ht) * (Ht(im + et0, jm + et1, km + et2) *
                                                      ut(im + et0, jm + et1, km + et2) -
                                                  Ht(im, jm, km) * ut(im, jm, km)) -
                                    dt3 * s(im, jm, km, comp) * dv(im, jm, km) +
                                    (dt3 / ht) * s(im, jm, km, comp) *
                                        (ut(im + et0, jm + et1, km + et2) - ut(im, jm, km));
                                r = R(i, j, k) -
                                    (dt3 / ht) * (Ht(i + et0, j + et1, k + et2) *
                                                      ut(i + et0, j + et1, k + et2) -
                                                  Ht(i, j, k) * ut(i, j, k)) -
                                    dt3 * s(i, j, k, comp) * dv(i, j, k) +
                                    (dt3 / ht) * s(i, j, k, comp) *
                                        (ut(i + et0, j + et1, k + et2) - ut(i, j, k));
                            } else {
                                l = L(i, j, k) -
                                    (dt6 / ht) *
                                        (ut(im + et0, jm + et1, km + et2) + ut(im, jm, km)) *
                                        (Ht(im + et0, jm + et1, km + et2) - Ht(im, jm, km));
                                r = R(i, j, k) -
                                    (dt6 / ht) * (ut(i + et0, j + et1, k + et2) + ut(i, j, k)) *
                                        (Ht(i + et0, j + et1, k + et2) - Ht(i, j, k));
                            }
                            const int pos = (d == 0) ? i : (d == 1) ? j : k;
                            if (pos == domlo[d]) {
                                if (bclo == MX_BC_EXT_DIR) {
                                    l = s(im, jm, km, comp);
                                    r = s(im, jm, km, comp);
                                } else if (bclo == MX_BC_FOEXTRAP || bclo == MX_BC_HOEXTRAP) {
                                    if (is_vel && comp == d) r = mn(r, 0.0);
                                    l = r;
                                } else if (bclo == MX_BC_REFLECT_EVEN) {
                                    l = r;
                                } else if (bclo == MX_BC_REFLECT_ODD) {
                                    l = 0.0;
                                    r = 0.0;
                                }
                            } else if (pos == domhi[d] + 1) {
                                if (bchi == MX_BC_EXT_DIR) {
                                    l = s(i, j, k, comp);
                                    r = s(i, j, k, comp);
                                } else if (bchi == MX_BC_FOEXTRAP || bchi == MX_BC_HOEXTRAP) {
                                    if (is_vel && comp == d) l = mx(l, 0.0);
                                    r = l;
                                } else if (bchi == MX_BC_REFLECT_EVEN) {
                                    r = l;
                                } else if (bchi == MX_BC_REFLECT_ODD) {
                                    l = 0.0;
                                    r = 0.0;
                                }
                            }
                            double res = (ud(i, j, k) > 0.0) ? l : r;
                            res = (ab(ud(i, j, k)) > rel_eps_local) ? res : 0.5 * (l + r);
                            out(i, j, k) = res;
                        });
                    }
                }
            }

            // ---- MakeEdgeScalEdges
            for (int d = 0; d < dim; ++d) {
                const IBox nbx = nodal(grids[b], d);
                const int e0 = e_(d, 0), e1 = e_(d, 1), e2 = e_(d, 2);
                const int bclo = bcs[bccomp].lo[d], bchi = bcs[bccomp].hi[d];
                const Arr& L = sl[d][b];
                const Arr& R = sr[d][b];
                const Arr& ud = umac[d][b];
                const Arr& out = sedge[d][b];
                const Arr& fo = force[b];
                const Arr& IpfA = Ipf[b];
                const Arr& ImfA = Imf[b];
                int tdir[2];
                int nt = 0;
                for (int t = 0; t < dim; ++t)
                    if (t != d) tdir[nt++] = t;
                const double hd = dx[d];
                par_for(nbx, [&](int i, int j, int k) {
                    const int im = i - e0, jm = j - e1, km = k - e2;
                    double fl = (ppm_trace_forces == 0) ? fo(im, jm, km, comp) : IpfA(im, jm, km, d);
                    double fr = (ppm_trace_forces == 0) ? fo(i, j, k, comp) : ImfA(i, j, k, d);
                    double l = L(i, j, k);
                    double r = R(i, j, k);
                    for (int q = 0; q < nt; ++q) {
                        const int t = tdir[q];
                        const int et0 = e_(t, 0), et1 = e_(t, 1), et2 = e_(t, 2);
                        const Arr& ut = umac[t][b];
                        const Arr& Tt = (dim == 3) ? T[t][3 - d - t][b] : simh[t][b];
                        const double ht = dx[t];
                        if (is_conservative) {
                            l = l - (dt2 / ht) * (Tt(im + et0, jm + et1, km + et2) *
                                                      ut(im + et0, jm + et1, km + et2) -
                                                  Tt(im, jm, km) * ut(im, jm, km));
                            r = r - (dt2 / ht) * (Tt(i + et0, j + et1, k + et2) *
                                                      ut(i + et0, j + et1, k + et2) -
                                                  Tt(i, j, k) * ut(i, j, k));
                        } else {
                            l = l - (dt4 / ht) *
                                        (ut(im + et0, jm + et1, km + et2) + ut(im, jm, km)) *
                                        (Tt(im + et0, jm + et1, km + et2) - Tt(im, jm, km));
                            r = r - (dt4 / ht) * (ut(i + et0, j + et1, k + et2) + ut(i, j, k)) *
                                        (Tt(i + et0, j + et1, k + et2) - Tt(i, j, k));
                        }
                    }
                    if (is_conservative) {
                        l = l - (dt2 / hd) * s(im, jm, km, comp) * (ud(i, j, k) - ud(im, jm, km));
                        r = r - (dt2 / hd) * s(i, j, k, comp) *
                                    (ud(i + e0, j + e1, k + e2) - ud(i, j, k));
                    }
                    double sedgel = l + dt2 * fl;
                    double sedger = r + dt2 * fr;

                    double res = (ud(i, j, k) > 0.0) ? sedgel : sedger;
                    res = (ab(ud(i, j, k)) > rel_eps_local) ? res : 0.5 * (sedgel + sedger);

                    const int pos = (d == 0) ? i : (d == 1) ? j : k;
                    if (pos == domlo[d]) {
                        if (bclo == MX_BC_EXT_DIR) {
                            res = s(im, jm, km, comp);
                        } else if (bclo == MX_BC_FOEXTRAP || bclo == MX_BC_HOEXTRAP) {
                            res = (is_vel && comp == d) ? mn(sedger, 0.0) : sedger;
                        } else if (bclo == MX_BC_REFLECT_EVEN) {
                            res = sedger;
                        } else if (bclo == MX_BC_REFLECT_ODD) {
                            res = 0.0;
                        }
                    } else if (pos == domhi[d] + 1) {
                        if (bchi == MX_BC_EXT_DIR) {
                            res = s(i, j, k, comp);
                        } else if (bchi == MX_BC_FOEXTRAP || bchi == MX_BC_HOEXTRAP) {
                            res = (is_vel && comp == d) ? mx(sedgel, 0.0) : sedgel;
                        } else if (bchi == MX_BC_REFLECT_EVEN) {
                            res = sedgel;
                        } else if (bchi == MX_BC_REFLECT_ODD) {
                            res = 0.0;
                        }
                    }
                    out(i, j, k, comp) = res;
                });
            }
        }
    }
    // AverageDownFaces(sedge) for is_vel is a no-op on a single level
    return MX_OK;
}

// ---------------------------------------------------------------------------
// Maestro::MakeRhoXFlux (MaestroMakeFlux.cpp:18-465), one level.
int MakeRhoXFlux(const Ctx& c, int lev, const std::vector<Arr> sflux[3],
                 const std::vector<Arr>& etarhoflux, const std::vector<Arr> sedge[3],
                 const std::vector<Arr> umac[3], const BaseArr& rho0_old,
                 const BaseArr& rho0_edge_old, const std::vector<Arr> r0mac_old[3],
                 const BaseArr& rho0_new, const BaseArr& rho0_edge_new,
                 const std::vector<Arr> r0mac_new[3], const BaseArr& rho0_predicted_edge,
                 const BaseArr& w0, int start_comp, int num_comp) {
    const int dim = c.dim();
    const int Rho = c.p.comp_rho;
    const int FirstSpec = c.p.first_spec;
    const int NumSpec = c.p.num_spec;
    const int species_pred_type = c.p.species_pred_type;
    const bool use_exact = c.p.use_exact_base_state != 0;
    const bool evolve = c.p.evolve_base_state != 0;
    const std::vector<IBox>& grids = c.grids[lev];

    for (int d = 0; d < dim; ++d)
        for (size_t b = 0; b < grids.size(); ++b) {
            const Arr& f = sflux[d][b];
            par_for(nodal(grids[b], d), [&](int i, int j, int k) { f(i, j, k, Rho) = 0.0; });
        }

    for (size_t b = 0; b < grids.size(); ++b) {
        for (int d = 0; d < dim; ++d) {
            const Arr& sf = sflux[d][b];
            const Arr& se = sedge[d][b];
            const Arr& um = umac[d][b];
            const bool vertical = (d == dim - 1);
            Arr eta;
            if (vertical) eta = etarhoflux[b];
            Arr r0o, r0n;
            if (dim == 3 && c.p.spherical) {
                r0o = r0mac_old[d][b];
                r0n = r0mac_new[d][b];
            }
            par_for(nodal(grids[b], d), [&](int i, int j, int k) {
                const int r = (dim == 2) ? j : k;
                for (int comp = start_comp; comp < num_comp + start_comp; ++comp) {
                    double rho0_edge;
                    if (dim == 3 && c.p.spherical) {
                        // LinComb(0.5, r0mac_old, 0.5, r0mac_new)
                        rho0_edge = 0.5 * r0o(i, j, k) + 0.5 * r0n(i, j, k);
                    } else if (vertical) {
                        rho0_edge = 0.5 * (rho0_edge_old(lev, r) + rho0_edge_new(lev, r));
                    } else {
                        rho0_edge = 0.5 * (rho0_old(lev, r) + rho0_new(lev, r));
                    }
                    if (species_pred_type == 1) {
                        sf(i, j, k, comp) = um(i, j, k) * (rho0_edge + se(i, j, k, Rho)) *
                                            se(i, j, k, comp);
                    } else if (species_pred_type == 2) {
                        sf(i, j, k, comp) = um(i, j, k) * se(i, j, k, comp);
                    } else if (species_pred_type == 3) {
                        sf(i, j, k, comp) = um(i, j, k) * se(i, j, k, Rho) * se(i, j, k, comp);
                    }
                    if (vertical && !(dim == 3 && c.p.spherical) && evolve && !use_exact) {
                        if (comp >= FirstSpec && comp <= FirstSpec + NumSpec - 1) {
                            eta(i, j, k) += sf(i, j, k, comp);
                        }
                        if (comp == FirstSpec + NumSpec - 1) {
                            eta(i, j, k) -= w0(lev, r) * rho0_predicted_edge(lev, r);
                        }
                    }
                    sf(i, j, k, Rho) += sf(i, j, k, comp);
                }
            });
        }
    }
    // reflux_type 1: AverageDownFaces(sflux) is a no-op on a single level;
    // reflux_type 2: the flux registers of lev and lev + 1
    return FluxRegOps(c, lev, sflux, true, start_comp, num_comp);
}

// Maestro::MakeRhoHFlux (MaestroMakeFlux.cpp:467-1067), one level.
int MakeRhoHFlux(const Ctx& c, int lev, const std::vector<Arr> sflux[3],
                 const std::vector<Arr> sedge[3], const std::vector<Arr> umac[3],
                 const BaseArr& rho0_old, const BaseArr& rho0_edge_old,
                 const std::vector<Arr> r0mac_old[3], const BaseArr& rho0_new,
                 const BaseArr& rho0_edge_new, const std::vector<Arr> r0mac_new[3],
                 const BaseArr& rhoh0_old, const BaseArr& rhoh0_edge_old,
                 const std::vector<Arr> rh0mac_old[3], const BaseArr& rhoh0_new,
                 const BaseArr& rhoh0_edge_new, const std::vector<Arr> rh0mac_new[3],
                 const std::vector<Arr> h0mac_old[3], const std::vector<Arr> h0mac_new[3]) {
    const int dim = c.dim();
    const int Rho = c.p.comp_rho;
    const int RhoH = c.p.comp_rhoh;
    const int ept = c.p.enthalpy_pred_type;
    const int spt = c.p.species_pred_type;
    const bool have_h = ept == 2 || ept == 4 || ept == 6;
    const bool have_hprime = ept == 5;
    const bool have_rhoh = ept == 0;
    const std::vector<IBox>& grids = c.grids[lev];
    const bool sph = (dim == 3 && c.p.spherical);
    const bool exact = c.p.use_exact_base_state != 0;

    for (size_t b = 0; b < grids.size(); ++b) {
        for (int d = 0; d < dim; ++d) {
            const Arr& sf = sflux[d][b];
            const Arr& se = sedge[d][b];
            const Arr& um = umac[d][b];
            const bool vertical = (d == dim - 1);
            Arr ro, rn, rho_, rhn, ho, hn;
            if (sph) {
                if (exact) {
                    rho_ = rh0mac_old[d][b];
                    rhn = rh0mac_new[d][b];
                } else {
                    ro = r0mac_old[d][b];
                    rn = r0mac_new[d][b];
                    ho = h0mac_old[d][b];
                    hn = h0mac_new[d][b];
                }
            }
            par_for(nodal(grids[b], d), [&](int i, int j, int k) {
                const int r = (dim == 2) ? j : k;
                if (!sph) {
                    if (have_h) {
                        if (spt == 1) {
                            double rho0_edge =
                                vertical ? 0.5 * (rho0_edge_old(lev, r) + rho0_edge_new(lev, r))
                                         : 0.5 * (rho0_old(lev, r) + rho0_new(lev, r));
                            sf(i, j, k, RhoH) = um(i, j, k) * (rho0_edge + se(i, j, k, Rho)) *
                                                se(i, j, k, RhoH);
                        } else if (spt == 3 || spt == 2) {
                            sf(i, j, k, RhoH) = um(i, j, k) * se(i, j, k, Rho) * se(i, j, k, RhoH);
                        }
                    } else if (have_rhoh) {
                        sf(i, j, k, RhoH) = um(i, j, k) * se(i, j, k, RhoH);
                    } else if (ept == 1 || ept == 3) {
                        double rhoh0_edge =
                            vertical ? 0.5 * (rhoh0_edge_old(lev, r) + rhoh0_edge_new(lev, r))
                                     : 0.5 * (rhoh0_old(lev, r) + rhoh0_new(lev, r));
                        sf(i, j, k, RhoH) = um(i, j, k) * (rhoh0_edge + se(i, j, k, RhoH));
                    }
                    // pred_hprime: no planar branch, the flux is left untouched
                } else if (exact) {
                    // rhoh0mac_edge = LinComb(0.5, rh0mac_old, 0.5, rh0mac_new)
                    double rhoh0_edge = 0.5 * rho_(i, j, k) + 0.5 * rhn(i, j, k);
                    if (d == 0) {
                        if (have_h || have_rhoh) {
                            // `if (have_h || have_rhoh) {}` shadows the have_rhoh branch
                        } else {
                            sf(i, j, k, RhoH) = um(i, j, k) * (rhoh0_edge + se(i, j, k, RhoH));
                        }
                    } else {
                        if (have_h || have_hprime) {
                        } else if (have_rhoh) {
                            sf(i, j, k, RhoH) = um(i, j, k) * se(i, j, k, RhoH);
                        } else {
                            sf(i, j, k, RhoH) = um(i, j, k) * (rhoh0_edge + se(i, j, k, RhoH));
                        }
                    }
                } else {
                    double rho0_edge = 0.5 * ro(i, j, k) + 0.5 * rn(i, j, k);
                    double h0_edge = 0.5 * ho(i, j, k) + 0.5 * hn(i, j, k);
                    if (have_h) {
                        if (spt == 1) {
                            sf(i, j, k, RhoH) = um(i, j, k) * (rho0_edge + se(i, j, k, Rho)) *
                                                se(i, j, k, RhoH);
                        } else if (spt == 3 || spt == 2) {
                            sf(i, j, k, RhoH) = um(i, j, k) * se(i, j, k, Rho) * se(i, j, k, RhoH);
                        }
                    } else if (have_hprime) {
                        if (spt == 1) {
                            sf(i, j, k, RhoH) = um(i, j, k) * (se(i, j, k, Rho) + rho0_edge) *
                                                (se(i, j, k, RhoH) + h0_edge);
                        }
                    } else if (have_rhoh) {
                        sf(i, j, k, RhoH) = um(i, j, k) * se(i, j, k, RhoH);
                    } else {
                        sf(i, j, k, RhoH) =
                            um(i, j, k) * (rho0_edge * h0_edge + se(i, j, k, RhoH));
                    }
                }
            });
        }
    }
    return FluxRegOps(c, lev, sflux, false, 0, 0);
}

// ---------------------------------------------------------------------------
// Maestro::CelltoEdge (MaestroCelltoEdge.cpp:5-55), planar, one radial level
// without coarse-fine chunks.
void CelltoEdge(int lev, int nr, const mx_basestate* cell, mx_basestate* edge) {
    auto s0_cell = [&](int r) { return cell->p[(long long)lev * cell->len + r]; };
    auto s0_edge = [&](int r) -> double& { return edge->p[(long long)lev * edge->len + r]; };
    const int lo = 0;
    const int hi = nr;
    const double nr_lev = nr;
    for (int r = lo; r <= hi; ++r) {
        if (r == 0) {
            s0_edge(r) = s0_cell(r);
        } else if (r == 1 || r == lo) {
            s0_edge(r) = 0.5 * (s0_cell(r - 1) + s0_cell(r));
        } else if (r == nr_lev) {
            s0_edge(r) = s0_cell(r - 1);
        } else if (r == nr_lev - 1 || r == hi) {
            s0_edge(r) = 0.5 * (s0_cell(r) + s0_cell(r - 1));
        } else {
            double tmp = 7.0 / 12.0 * (s0_cell(r) + s0_cell(r - 1)) -
                         1.0 / 12.0 * (s0_cell(r + 1) + s0_cell(r - 2));
            double s0min = mn(s0_cell(r), s0_cell(r - 1));
            double s0max = mx(s0_cell(r), s0_cell(r - 1));
            s0_edge(r) = mn(mx(tmp, s0min), s0max);
        }
    }
}

}  // namespace mxo
