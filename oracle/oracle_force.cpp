// oracle_force.cpp -- CPU restatement of the force producers of the pass:
// Maestro::MakeVelForce and Maestro::ModifyScalForce, one level each.
//
// TEST INFRASTRUCTURE ONLY (see oracle_core.H).  Loop bodies keep the
// reference's operation order.  The Cartesian base-state arrays the reference
// builds with Put1dArrayOnCart (rho0, grav_cell, gradw0, divw0, s0_edge) are
// evaluated per cell by the same formulas (planar row copy; spherical
// non-exact interpolation).  ROTATION (a compile-time option of the
// reference, off by default) is not restated.
#include "oracle_core.H"

namespace mxo {

int Put1dArrayOnCart(const Ctx& c, int lev, const BaseArr& s0, const std::vector<Arr>& s0_cart,
                     bool edge, bool vec, const mx_sph_geom* sg);

namespace {

// a temporary single-level MultiFab over each box grown by `ng` cells
struct TmpLevel {
    std::vector<std::vector<double>> store;
    std::vector<Arr> a;
    TmpLevel(const std::vector<IBox>& grids, int ncomp, int ng, int dim) {
        store.resize(grids.size());
        a.resize(grids.size());
        for (size_t b = 0; b < grids.size(); ++b) {
            IBox g = grow(grids[b], dim, ng);
            const long long n0 = g.hi[0] - g.lo[0] + 1, n1 = g.hi[1] - g.lo[1] + 1,
                            n2 = g.hi[2] - g.lo[2] + 1;
            store[b].assign(n0 * n1 * n2 * ncomp, 0.0);
            Arr& x = a[b];
            x.p = store[b].data();
            x.js = n0;
            x.ks = n0 * n1;
            x.ns = n0 * n1 * n2;
            for (int d = 0; d < 3; ++d) {
                x.b[d] = g.lo[d];
                x.e[d] = g.hi[d] + 1;
            }
            x.nc = ncomp;
        }
    }
};

// a level-wide Put1dArrayOnCart into a temporary of the valid cells
int cart_of(const Ctx& c, int lev, const BaseArr& s0, bool vec, const mx_sph_geom* sg,
            TmpLevel& out) {
    return Put1dArrayOnCart(c, lev, s0, out.a, false, vec, sg);
}

}  // namespace

// Maestro::MakeVelForce on one level (MaestroForce.cpp:6-258), without
// ROTATION: vel_force on the valid cells (the reference zeroes the FAB first
// and FillPatch then fills the ghosts; ghosts are the caller's here).
// rho: component 0 of the state; gpi, w0_force_cart, w0_cart: dim
// components; normal: 3 components (spherical).  gradw0 is zero when the
// base state is exact or averaged, as in the reference.
int MakeVelForce(const Ctx& c, int lev, const std::vector<Arr>& vel_force,
                 const std::vector<Arr> uedge[3], const std::vector<Arr>& rho,
                 const std::vector<Arr>& gpi, const std::vector<Arr>& w0_force_cart,
                 const std::vector<Arr>& w0_cart, const std::vector<Arr>& normal,
                 const BaseArr& rho0, const BaseArr& grav_cell, const BaseArr& w0,
                 double buoyancy_cutoff_factor, double base_cutoff_density,
                 bool do_add_utilde_force, bool average_base_state, const mx_sph_geom* sg) {
    const int dim = c.dim();
    const bool sph = c.p.spherical != 0;
    const std::vector<IBox>& grids = c.grids[lev];
    const double* dx = c.p.dx[lev];
    const int domhi = c.domain(lev).hi[dim - 1];
    if (sph && !sg) return MX_ERR_ARG;
    // gradw0 (MaestroForce.cpp:34-44): spherical uses row 0 over nr_fine
    std::vector<double> gw;
    BaseArr gradw0;
    if (sph) {
        const int nr = sg->nr_fine;
        gw.assign(nr, 0.0);
        if (!(c.p.use_exact_base_state || average_base_state))
            for (int r = 0; r < nr; ++r) gw[r] = (w0(0, r + 1) - w0(0, r)) / sg->dr_fine;
        gradw0.p = gw.data();
        gradw0.len = nr;
    }
    TmpLevel rho0_cart(grids, 1, 0, dim), grav_cart(grids, 3, 0, dim), gradw0_cart(grids, 1, 0, dim);
    int rc = cart_of(c, lev, rho0, false, sg, rho0_cart);
    if (!rc) rc = cart_of(c, lev, grav_cell, true, sg, grav_cart);
    if (!rc && sph) rc = cart_of(c, lev, gradw0, false, sg, gradw0_cart);
    if (rc) return rc;
    const double cutoff = buoyancy_cutoff_factor * base_cutoff_density;
    for (size_t b = 0; b < grids.size(); ++b) {
        const Arr& vf = vel_force[b];
        const Arr& ro = rho[b];
        const Arr& gp = gpi[b];
        const Arr& wf = w0_force_cart[b];
        const Arr& gr = grav_cart.a[b];
        const Arr& r0 = rho0_cart.a[b];
        // the edge velocities and w0_cart are read by the Utilde term only
        Arr ue, ve, we, w0c, nrm;
        if (do_add_utilde_force) {
            ue = uedge[0][b];
            ve = uedge[1][b];
            if (dim == 3) we = uedge[2][b];
            if (!sph) w0c = w0_cart[b];
        }
        if (sph) nrm = normal[b];
        if (!sph) {
            par_for(grids[b], [&](int i, int j, int k) {
                double rhopert = ro(i, j, k) - r0(i, j, k);
                if (ro(i, j, k) < cutoff) rhopert = 0.0;
                for (int d = 0; d < dim - 1; ++d) vf(i, j, k, d) = -gp(i, j, k, d) / ro(i, j, k);
                vf(i, j, k, dim - 1) =
                    (rhopert * gr(i, j, k, dim - 1) - gp(i, j, k, dim - 1)) / ro(i, j, k) -
                    wf(i, j, k, dim - 1);
                if (do_add_utilde_force) {
                    if (dim == 2) {
                        if (!(j <= -1 || j >= domhi))
                            vf(i, j, k, 1) -= (ve(i, j + 1, k) + ve(i, j, k)) *
                                              (w0c(i, j + 1, k, 1) - w0c(i, j, k, 1)) /
                                              (2.0 * dx[1]);
                    } else {
                        if (!(k <= -1 || k >= domhi))
                            vf(i, j, k, 2) -= (we(i, j, k + 1) + we(i, j, k)) *
                                              (w0c(i, j, k + 1, 2) - w0c(i, j, k, 2)) /
                                              (2.0 * dx[2]);
                    }
                }
            });
        } else {
            const Arr& gw0 = gradw0_cart.a[b];
            par_for(grids[b], [&](int i, int j, int k) {
                double rhopert = ro(i, j, k) - r0(i, j, k);
                if (ro(i, j, k) < cutoff) rhopert = 0.0;
                for (int d = 0; d < 3; ++d)
                    vf(i, j, k, d) = (rhopert * gr(i, j, k, d) - gp(i, j, k, d)) / ro(i, j, k) -
                                     wf(i, j, k, d);
                if (do_add_utilde_force) {
                    const double ut = 0.5 * (ue(i, j, k) + ue(i + 1, j, k)) * nrm(i, j, k, 0) +
                                      0.5 * (ve(i, j, k) + ve(i, j + 1, k)) * nrm(i, j, k, 1) +
                                      0.5 * (we(i, j, k) + we(i, j, k + 1)) * nrm(i, j, k, 2);
                    for (int d = 0; d < 3; ++d)
                        vf(i, j, k, d) -= ut * gw0(i, j, k) * nrm(i, j, k, d);
                }
            });
        }
    }
    return MX_OK;
}

// Maestro::ModifyScalForce on one level (MaestroForce.cpp:261-466),
// component comp of scal_force on the valid cells.  Planar: s0_edge put on
// the cells (the edge below each cell); the cell above a box's top cell takes
// the same formula inside the domain and, past a non-periodic domain top,
// the value of the top cell (first-order extrapolation, the FillPatch of
// the forcing boundary conditions at level 0), the periodic image past a
// periodic one.  Spherical: divw0 from w0 (non-exact base state) put on the
// cells.  s0_cart: 1 component with 1 ghost cell; w0_cart: dim components
// with 1 ghost cell (planar).
int ModifyScalForce(const Ctx& c, int lev, const std::vector<Arr>& scal_force,
                    const std::vector<Arr>& state, const std::vector<Arr> umac[3],
                    const BaseArr& s0_edge, const std::vector<Arr>& s0_cart,
                    const std::vector<Arr>& w0_cart, const BaseArr& w0, int comp, bool fullform,
                    const mx_sph_geom* sg) {
    const int dim = c.dim();
    const bool sph = c.p.spherical != 0;
    const std::vector<IBox>& grids = c.grids[lev];
    const double* dx = c.p.dx[lev];
    const IBox dom = c.domain(lev);
    if (sph && !sg) return MX_ERR_ARG;
    TmpLevel divu_cart(grids, 1, 0, dim);
    std::vector<double> dw;
    if (sph) {
        const int nr = sg->nr_fine;
        dw.assign(nr, 0.0);
        if (!c.p.use_exact_base_state) {
            const double* re = sg->r_edge_loc;
            const double* rc = sg->r_cc_loc;
            for (int r = 0; r < nr - 1; ++r) {
                const double dr_loc = re[r + 1] - re[r];
                dw[r] = (re[r + 1] * re[r + 1] * w0(0, r + 1) - re[r] * re[r] * w0(0, r)) /
                        (dr_loc * rc[r] * rc[r]);
            }
        }
        BaseArr divw0;
        divw0.p = dw.data();
        divw0.len = nr;
        int rc = cart_of(c, lev, divw0, false, sg, divu_cart);
        if (rc) return rc;
    }
    const int vert = dim - 1;
    const bool vper = c.p.is_periodic[vert] != 0;
    const int vlo = dom.lo[vert], vhi = dom.hi[vert];
    // s0_edge put on cell r: the edge below it (the planar s0_edge_cart)
    auto s0e = [&](int r) {
        if (r > vhi) r = vper ? r - (vhi - vlo + 1) : vhi;
        return s0_edge(lev, r);
    };
    for (size_t b = 0; b < grids.size(); ++b) {
        const Arr& f = scal_force[b];
        const Arr& s = state[b];
        const Arr& um = umac[0][b];
        const Arr& vm = umac[1][b];
        const Arr& s0 = s0_cart[b];
        Arr wm, w0c;
        if (dim == 3) wm = umac[2][b];
        if (!sph) w0c = w0_cart[b];
        if (sph) {
            const Arr& dv = divu_cart.a[b];
            par_for(grids[b], [&](int i, int j, int k) {
                const double divumac = (um(i + 1, j, k) - um(i, j, k)) / dx[0] +
                                       (vm(i, j + 1, k) - vm(i, j, k)) / dx[1] +
                                       (wm(i, j, k + 1) - wm(i, j, k)) / dx[2];
                if (fullform) {
                    f(i, j, k, comp) -= s(i, j, k, comp) * (divumac + dv(i, j, k));
                } else {
                    const double c0 = s0(i, j, k);
                    const double xhi = i < dom.hi[0] ? 0.5 * (c0 + s0(i + 1, j, k)) : c0;
                    const double xlo = i > dom.lo[0] ? 0.5 * (c0 + s0(i - 1, j, k)) : c0;
                    const double yhi = j < dom.hi[1] ? 0.5 * (c0 + s0(i, j + 1, k)) : c0;
                    const double ylo = j > dom.lo[1] ? 0.5 * (c0 + s0(i, j - 1, k)) : c0;
                    const double zhi = k < dom.hi[2] ? 0.5 * (c0 + s0(i, j, k + 1)) : c0;
                    const double zlo = k > dom.lo[2] ? 0.5 * (c0 + s0(i, j, k - 1)) : c0;
                    const double divs0u = (um(i + 1, j, k) * xhi - um(i, j, k) * xlo) / dx[0] +
                                          (vm(i, j + 1, k) * yhi - vm(i, j, k) * ylo) / dx[1] +
                                          (wm(i, j, k + 1) * zhi - wm(i, j, k) * zlo) / dx[2];
                    f(i, j, k, comp) +=
                        -divs0u - (s(i, j, k, comp) - c0) * (divumac + dv(i, j, k));
                }
            });
        } else {
            par_for(grids[b], [&](int i, int j, int k) {
                double divu = (um(i + 1, j, k) - um(i, j, k)) / dx[0] +
                              (vm(i, j + 1, k) - vm(i, j, k)) / dx[1];
                if (dim == 3) divu = divu + (wm(i, j, k + 1) - wm(i, j, k)) / dx[2];
                if (dim == 2)
                    divu += (w0c(i, j + 1, k, 1) - w0c(i, j, k, 1)) / dx[1];
                else
                    divu += (w0c(i, j, k + 1, 2) - w0c(i, j, k, 2)) / dx[2];
                if (fullform) {
                    f(i, j, k, comp) -= s(i, j, k, comp) * divu;
                } else {
                    double divs0u;
                    if (dim == 2)
                        divs0u = s0(i, j, k) * (um(i + 1, j, k) - um(i, j, k)) / dx[0] +
                                 (vm(i, j + 1, k) * s0e(j + 1) - vm(i, j, k) * s0e(j)) / dx[1];
                    else
                        divs0u = s0(i, j, k) * ((um(i + 1, j, k) - um(i, j, k)) / dx[0] +
                                                (vm(i, j + 1, k) - vm(i, j, k)) / dx[1]) +
                                 (wm(i, j, k + 1) * s0e(k + 1) - wm(i, j, k) * s0e(k)) / dx[2];
                    f(i, j, k, comp) += -(s(i, j, k, comp) - s0(i, j, k)) * divu - divs0u;
                }
            });
        }
    }
    return MX_OK;
}

}  // namespace mxo
