// oracle_base.cpp -- CPU restatement of the base-state <-> Cartesian producers
// that feed the advection pass: Put1dArrayOnCart, MakeW0mac, MakeS0mac,
// PutInPertForm, ConvertRhoXToX and ConvertRhoHToH (one level each).
//
// TEST INFRASTRUCTURE ONLY (see oracle_core.H).  The loop bodies keep the
// reference's operation order.  The reference's multi-level wrappers follow
// each level with AverageDown and FillPatch; like the library, the oracle
// leaves those to the caller.  Spherical geometry is the non-exact base state
// (use_exact_base_state = 0), whose radial index is int(radius / dr_fine).
#include "oracle_core.H"

namespace mxo {

namespace {

// a base-state read; the bounds-checked build records a miss
double at(const BaseArr& a, int lev, int r) {
#ifdef MXO_BOUNDS
    if (r < 0 || r >= a.len) {
        g_oob.store(1);
        return 0.0;
    }
#endif
    return a(lev, r);
}

double radial(const double* v, int n, int r) {
#ifdef MXO_BOUNDS
    if (r < 0 || r >= n) {
        g_oob.store(1);
        return 0.0;
    }
#endif
    (void)n;
    return v[r];
}

// QuadInterp (MaestroFill3dData.cpp:341-356) with limit = true, the default
// of its declaration (Maestro.H:63-68)
double quad_interp(double x, double x0, double x1, double x2, double y0, double y1, double y2) {
    double y = y0 + (y1 - y0) / (x1 - x0) * (x - x0) +
               ((y2 - y1) / (x2 - x1) - (y1 - y0) / (x1 - x0)) / (x2 - x0) * (x - x0) * (x - x1);
    if (y > mx(y0, mx(y1, y2))) y = mx(y0, mx(y1, y2));
    if (y < mn(y0, mn(y1, y2))) y = mn(y0, mn(y1, y2));
    return y;
}

struct Sph {
    const mx_sph_geom& g;
    const double* dx;
    int ncc() const { return g.nr_fine; }
    double rcc(int r) const { return radial(g.r_cc_loc, g.nr_fine, r); }
    double redge(int r) const { return radial(g.r_edge_loc, g.nr_fine + 1, r); }
    // coordinates of a point relative to the centre: cell centre (h = 0.5)
    // or face / node (h = 0) in each direction
    void pos(int i, int j, int k, double h0, double h1, double h2, double& x, double& y,
             double& z) const {
        x = g.prob_lo[0] + (double(i) + h0) * dx[0] - g.center[0];
        y = g.prob_lo[1] + (double(j) + h1) * dx[1] - g.center[1];
        z = g.prob_lo[2] + (double(k) + h2) * dx[2] - g.center[2];
    }
    void pos_face(int d, int i, int j, int k, double& x, double& y, double& z) const {
        pos(i, j, k, d == 0 ? 0.0 : 0.5, d == 1 ? 0.0 : 0.5, d == 2 ? 0.0 : 0.5, x, y, z);
    }
};

// s0 of an edge-centred base state at radius (w0_interp_type 1-3,
// MaestroFill3dData.cpp:211-253); non-exact base state
double edge_value(const Sph& S, const BaseArr& s0, double radius, int type) {
    const double drf = S.g.dr_fine;
    const int nr_fine = S.g.nr_fine;
    int index = int(radius / drf);
    const double rfac = (radius - double(index) * drf) / drf;
    double v = 0.0;
    if (type == 1) {
        v = rfac > 0.5 ? at(s0, 0, index + 1) : at(s0, 0, index);
    } else if (type == 2) {
        if (index < nr_fine)
            v = rfac * at(s0, 0, index + 1) + (1.0 - rfac) * at(s0, 0, index);
        else
            v = at(s0, 0, nr_fine);
    } else if (type == 3) {
        if (index <= 0)
            index = 0;
        else if (index >= nr_fine - 1)
            index = nr_fine - 2;
        else if (radius - S.redge(index) < S.redge(index + 1))
            index--;
        v = quad_interp(radius, S.redge(index), S.redge(index + 1), S.redge(index + 2),
                        at(s0, 0, index), at(s0, 0, index + 1), at(s0, 0, index + 2));
    }
    return v;
}

// s0 of a cell-centred base state at radius (s0_interp_type 1-3,
// MaestroFill3dData.cpp:271-325); non-exact base state
double cell_value(const Sph& S, const BaseArr& s0, double radius, int type) {
    const double drf = S.g.dr_fine;
    const int nr_fine = S.g.nr_fine;
    int index = int(radius / drf);
    double v = 0.0;
    if (type == 1) {
        v = at(s0, 0, index);
    } else if (type == 2) {
        if (radius >= S.rcc(index)) {
            if (index >= nr_fine - 1)
                v = at(s0, 0, nr_fine - 1);
            else
                v = at(s0, 0, index + 1) * (radius - S.rcc(index)) / drf +
                    at(s0, 0, index) * (S.rcc(index + 1) - radius) / drf;
        } else {
            if (index == 0)
                v = at(s0, 0, index);
            else if (index > nr_fine - 1)
                v = at(s0, 0, nr_fine - 1);
            else
                v = at(s0, 0, index) * (radius - S.rcc(index - 1)) / drf +
                    at(s0, 0, index - 1) * (S.rcc(index) - radius) / drf;
        }
    } else if (type == 3) {
        if (index == 0)
            index = 1;
        else if (index >= nr_fine - 1)
            index = nr_fine - 2;
        v = quad_interp(radius, S.rcc(index - 1), S.rcc(index), S.rcc(index + 1),
                        at(s0, 0, index - 1), at(s0, 0, index), at(s0, 0, index + 1));
    }
    return v;
}

// s0mac of a cell-centred base state on a face at radius (s0mac_interp_type
// 2, 3; MaestroFill3dData.cpp:1009-1188): the same formulas as cell_value
// with types 2 and 3
double face_value(const Sph& S, const BaseArr& s0, double radius, int type) {
    return cell_value(S, s0, radius, type);
}

}  // namespace

// Maestro::Put1dArrayOnCart on one level (MaestroFill3dData.cpp:49-339),
// valid cells; s0_cart gets component dim - 1 (planar vector), 0 (scalar) or
// 0..2 (spherical vector)
int Put1dArrayOnCart(const Ctx& c, int lev, const BaseArr& s0, const std::vector<Arr>& s0_cart,
                     bool edge, bool vec, const mx_sph_geom* sg) {
    const int dim = c.dim();
    const std::vector<IBox>& grids = c.grids[lev];
    const bool sph = c.p.spherical != 0;
    if (sph && (!sg || c.p.use_exact_base_state)) return MX_ERR_UNSUPPORTED;
    for (size_t b = 0; b < grids.size(); ++b) {
        const Arr& sc = s0_cart[b];
        if (!sph) {
            const int outcomp = vec ? dim - 1 : 0;
            par_for(grids[b], [&](int i, int j, int k) {
                const int r = dim == 2 ? j : k;
                sc(i, j, k, outcomp) =
                    edge ? 0.5 * (at(s0, lev, r) + at(s0, lev, r + 1)) : at(s0, lev, r);
            });
        } else {
            const Sph S{*sg, c.p.dx[lev]};
            par_for(grids[b], [&](int i, int j, int k) {
                double x, y, z;
                S.pos(i, j, k, 0.5, 0.5, 0.5, x, y, z);
                const double radius = std::sqrt(x * x + y * y + z * z);
                const double v = edge ? edge_value(S, s0, radius, sg->w0_interp_type)
                                      : cell_value(S, s0, radius, sg->s0_interp_type);
                if (vec) {
                    sc(i, j, k, 0) = v * x / radius;
                    sc(i, j, k, 1) = v * y / radius;
                    sc(i, j, k, 2) = v * z / radius;
                } else {
                    sc(i, j, k, 0) = v;
                }
            });
        }
    }
    return MX_OK;
}

// Maestro::MakeW0mac on one level (MaestroFill3dData.cpp:464-710): w0mac on
// every face of each box grown by one cell
int MakeW0mac(const Ctx& c, int lev, const BaseArr& w0, const std::vector<Arr>& w0_cart,
              const std::vector<Arr> w0mac[3], const mx_sph_geom* sg) {
    if (!c.p.spherical || c.dim() != 3 || !sg || c.p.use_exact_base_state)
        return MX_ERR_UNSUPPORTED;
    const int type = sg->w0mac_interp_type;
    if (type < 1 || type > 4) return MX_ERR_ARG;
    const std::vector<IBox>& grids = c.grids[lev];
    const Sph S{*sg, c.p.dx[lev]};
    const double drf = sg->dr_fine;
    const int nr_fine = sg->nr_fine;
    // type 4: w0 at the nodes, component n
    auto node_w0 = [&](int i, int j, int k, int n) {
        double x, y, z;
        S.pos(i, j, k, 0.0, 0.0, 0.0, x, y, z);
        const double radius = std::sqrt(x * x + y * y + z * z);
        const int index = int(radius / drf);
        const double rfac = (radius - double(index) * drf) / drf;
        double v;
        if (index < nr_fine)
            v = rfac * at(w0, 0, index + 1) + (1.0 - rfac) * at(w0, 0, index);
        else
            v = at(w0, 0, nr_fine);
        if (radius == 0.0) return v;
        const double q = n == 0 ? x : (n == 1 ? y : z);
        return v * q / radius;
    };
    for (size_t b = 0; b < grids.size(); ++b) {
        for (int d = 0; d < 3; ++d) {
            const Arr& wm = w0mac[d][b];
            IBox fb = grow(nodal(grids[b], d), 3, 1);
            if (type == 1) {
                const Arr& wc = w0_cart[b];
                par_for(fb, [&](int i, int j, int k) {
                    wm(i, j, k) = 0.5 * (wc(i - (d == 0), j - (d == 1), k - (d == 2), d) +
                                         wc(i, j, k, d));
                });
            } else if (type == 2 || type == 3) {
                par_for(fb, [&](int i, int j, int k) {
                    double x, y, z;
                    S.pos_face(d, i, j, k, x, y, z);
                    const double radius = std::sqrt(x * x + y * y + z * z);
                    const double v = edge_value(S, w0, radius, type);
                    const double q = d == 0 ? x : (d == 1 ? y : z);
                    wm(i, j, k) = v * q / radius;
                });
            } else {
                par_for(fb, [&](int i, int j, int k) {
                    // the four nodes of the face, in the reference's order
                    const int a0 = d == 0 ? 1 : 0;       // first transverse direction
                    const int a1 = d == 2 ? 1 : 2;       // second
                    int p1[3] = {i, j, k}, p2[3] = {i, j, k}, p3[3] = {i, j, k};
                    p1[a0] += 1;
                    p2[a1] += 1;
                    p3[a0] += 1;
                    p3[a1] += 1;
                    wm(i, j, k) = 0.25 * (node_w0(i, j, k, d) + node_w0(p1[0], p1[1], p1[2], d) +
                                          node_w0(p2[0], p2[1], p2[2], d) +
                                          node_w0(p3[0], p3[1], p3[2], d));
                });
            }
        }
    }
    return MX_OK;
}

// Maestro::MakeS0mac on one level (MaestroFill3dData.cpp:712-1188), the
// non-exact base state: s0mac on every face of each box grown by one cell.
// Type 1 averages s0_cart, which the reference builds with Put1dArrayOnCart
// and fills with FillPatch; here the caller passes it with 2 ghost cells.
int MakeS0mac(const Ctx& c, int lev, const BaseArr& s0, const std::vector<Arr>& s0_cart,
              const std::vector<Arr> s0mac[3], const mx_sph_geom* sg) {
    if (!c.p.spherical || c.dim() != 3 || !sg || c.p.use_exact_base_state)
        return MX_ERR_UNSUPPORTED;
    const int type = sg->s0mac_interp_type;
    if (type < 1 || type > 3) return MX_ERR_ARG;
    const std::vector<IBox>& grids = c.grids[lev];
    const Sph S{*sg, c.p.dx[lev]};
    for (size_t b = 0; b < grids.size(); ++b) {
        for (int d = 0; d < 3; ++d) {
            const Arr& sm = s0mac[d][b];
            IBox fb = grow(nodal(grids[b], d), 3, 1);
            if (type == 1) {
                const Arr& sc = s0_cart[b];
                par_for(fb, [&](int i, int j, int k) {
                    sm(i, j, k) = 0.5 * (sc(i - (d == 0), j - (d == 1), k - (d == 2)) + sc(i, j, k));
                });
            } else {
                par_for(fb, [&](int i, int j, int k) {
                    double x, y, z;
                    S.pos_face(d, i, j, k, x, y, z);
                    const double radius = std::sqrt(x * x + y * y + z * z);
                    sm(i, j, k) = face_value(S, s0, radius, type);
                });
            }
        }
    }
    return MX_OK;
}

// Maestro::PutInPertForm on one level (MaestroConvert.cpp:36-62): scal[comp]
// -= (flag) or += s0 put on the cells (cell-centred, scalar), valid cells
int PutInPertForm(const Ctx& c, int lev, const std::vector<Arr>& scal, const BaseArr& s0,
                  int comp, bool flag, const mx_sph_geom* sg) {
    const std::vector<IBox>& grids = c.grids[lev];
    std::vector<std::vector<double>> store(grids.size());
    std::vector<Arr> s0_cart(grids.size());
    for (size_t b = 0; b < grids.size(); ++b) {
        const IBox& g = grids[b];
        const long long n0 = g.hi[0] - g.lo[0] + 1, n1 = g.hi[1] - g.lo[1] + 1,
                        n2 = g.hi[2] - g.lo[2] + 1;
        store[b].assign(n0 * n1 * n2, 0.0);
        Arr& a = s0_cart[b];
        a.p = store[b].data();
        a.js = n0;
        a.ks = n0 * n1;
        a.ns = n0 * n1 * n2;
        for (int d = 0; d < 3; ++d) {
            a.b[d] = g.lo[d];
            a.e[d] = g.hi[d] + 1;
        }
        a.nc = 1;
    }
    int rc = Put1dArrayOnCart(c, lev, s0, s0_cart, false, false, sg);
    if (rc) return rc;
    for (size_t b = 0; b < grids.size(); ++b) {
        const Arr& s = scal[b];
        const Arr& sc = s0_cart[b];
        par_for(grids[b], [&](int i, int j, int k) {
            if (flag)
                s(i, j, k, comp) -= sc(i, j, k);
            else
                s(i, j, k, comp) += sc(i, j, k);
        });
    }
    return MX_OK;
}

// Maestro::ConvertRhoXToX (MaestroConvert.cpp:64-91) and ConvertRhoHToH
// (:93-114) on one level, valid cells: flag divides by Rho, else multiplies
int ConvertByRho(const Ctx& c, int lev, const std::vector<Arr>& scal, int comp0, int ncomp,
                 bool flag) {
    const std::vector<IBox>& grids = c.grids[lev];
    const int Rho = c.p.comp_rho;
    for (size_t b = 0; b < grids.size(); ++b) {
        const Arr& s = scal[b];
        for (int comp = comp0; comp < comp0 + ncomp; ++comp) {
            par_for(grids[b], [&](int i, int j, int k) {
                if (flag)
                    s(i, j, k, comp) /= s(i, j, k, Rho);
                else
                    s(i, j, k, comp) *= s(i, j, k, Rho);
            });
        }
    }
    return MX_OK;
}

}  // namespace mxo
