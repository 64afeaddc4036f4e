// oracle_update.cpp -- CPU restatement of Maestro::UpdateScal and
// Maestro::UpdateVel, the conservative update that closes the advection step.
//
// TEST INFRASTRUCTURE ONLY (see oracle_core.H).  Each loop body keeps the
// reference's operation order, so that the HIP kernels can be compared with
// it bit for bit.  The reference's reflux, AverageDown and FillPatch after
// the loops are multi-level synchronisation; like the library, the oracle
// leaves them to the caller.
#include "oracle_core.H"

namespace mxo {

// Maestro::UpdateScal (MaestroAdvection.cpp:86-297), the ParallelFor bodies
// on the valid cells of one level.  start_comp == RhoH: the enthalpy update
// (:120-162); start_comp == FirstSpec: the species update (:164-183) and the
// density, auxiliary, floor and negative-species pass (:185-249).
int UpdateScal(const Ctx& c, int lev, const std::vector<Arr>& sold, const std::vector<Arr>& snew,
               const std::vector<Arr> sflux[3], const std::vector<Arr>& force, int start_comp,
               double base_cutoff_density) {
    const int dim = c.dim();
    const double dt_loc = c.p.dt;
    const double* dx = c.p.dx[lev];
    const int Rho = c.p.comp_rho, RhoH = c.p.comp_rhoh;
    const int FirstSpec = c.p.first_spec, NumSpec = c.p.num_spec;
    const int FirstAux = FirstSpec + NumSpec;
    const int NumAux = c.p.nscal - FirstAux;
    const std::vector<IBox>& grids = c.grids[lev];
    for (size_t b = 0; b < grids.size(); ++b) {
        const Arr& so = sold[b];
        const Arr& sn = snew[b];
        const Arr& fx = sflux[0][b];
        const Arr& fy = sflux[1][b];
        const Arr& fo = force[b];
        Arr fz;
        if (dim == 3) fz = sflux[2][b];
        if (start_comp == RhoH) {
            par_for(grids[b], [&](int i, int j, int k) {
                double divterm = (fx(i + 1, j, k, RhoH) - fx(i, j, k, RhoH)) / dx[0] +
                                 (fy(i, j + 1, k, RhoH) - fy(i, j, k, RhoH)) / dx[1];
                if (dim == 3) divterm = divterm + (fz(i, j, k + 1, RhoH) - fz(i, j, k, RhoH)) / dx[2];
                sn(i, j, k, RhoH) = so(i, j, k, RhoH) + dt_loc * (-divterm + fo(i, j, k, RhoH));
            });
        } else if (start_comp == FirstSpec) {
            par_for(grids[b], [&](int i, int j, int k) {
                for (int n = 0; n < NumSpec; ++n) {
                    const int comp = FirstSpec + n;
                    double divterm = (fx(i + 1, j, k, comp) - fx(i, j, k, comp)) / dx[0];
                    divterm += (fy(i, j + 1, k, comp) - fy(i, j, k, comp)) / dx[1];
                    if (dim == 3) divterm += (fz(i, j, k + 1, comp) - fz(i, j, k, comp)) / dx[2];
                    sn(i, j, k, comp) = so(i, j, k, comp) + dt_loc * (-divterm + fo(i, j, k, comp));
                }
            });
            par_for(grids[b], [&](int i, int j, int k) {
                sn(i, j, k, Rho) = so(i, j, k, Rho);
                bool has_negative_species = false;
                for (int comp = start_comp; comp < start_comp + NumSpec; ++comp) {
                    sn(i, j, k, Rho) += sn(i, j, k, comp) - so(i, j, k, comp);
                    if (sn(i, j, k, comp) < 0.0) has_negative_species = true;
                }
                for (int comp = FirstAux; comp < FirstAux + NumAux; ++comp)
                    sn(i, j, k, comp) = so(i, j, k, comp) * sn(i, j, k, Rho) / so(i, j, k, Rho);
                if (sn(i, j, k, Rho) < 0.5 * base_cutoff_density) {
                    for (int comp = start_comp; comp < start_comp + NumSpec; ++comp)
                        sn(i, j, k, comp) *= 0.5 * base_cutoff_density / sn(i, j, k, Rho);
                    sn(i, j, k, Rho) = 0.5 * base_cutoff_density;
                }
                if (has_negative_species) {
                    for (int comp = start_comp; comp < start_comp + NumSpec; ++comp) {
                        if (sn(i, j, k, comp) < 0.0) {
                            const double delta = -sn(i, j, k, comp);
                            double sumX = 0.0;
                            for (int comp2 = start_comp; comp2 < start_comp + NumSpec; ++comp2)
                                if (comp2 != comp && sn(i, j, k, comp2) >= 0.0)
                                    sumX += sn(i, j, k, comp2);
                            for (int comp2 = start_comp; comp2 < start_comp + NumSpec; ++comp2) {
                                if (comp2 != comp && sn(i, j, k, comp2) >= 0.0) {
                                    const double frac = sn(i, j, k, comp2) / sumX;
                                    sn(i, j, k, comp2) -= frac * delta;
                                }
                            }
                            sn(i, j, k, comp) = 0.0;
                        }
                    }
                }
            });
        } else {
            return MX_ERR_ARG;
        }
    }
    return MX_OK;
}

// Maestro::UpdateVel (MaestroAdvection.cpp:299-526) on the valid cells of one
// level: planar (:335-422) or spherical (:423-505).
int UpdateVel(const Ctx& c, int lev, const std::vector<Arr>& uold, const std::vector<Arr>& unew,
              const std::vector<Arr> umac[3], const std::vector<Arr> uedge[3],
              const std::vector<Arr>& force, const std::vector<Arr>& sponge,
              const std::vector<Arr> w0mac[3], const std::vector<Arr>& w0_cart, bool do_sponge) {
    const int dim = c.dim();
    const double dt_loc = c.p.dt;
    const double* dx = c.p.dx[lev];
    const std::vector<IBox>& grids = c.grids[lev];
    if (c.p.spherical && dim != 3) return MX_ERR_UNSUPPORTED;
    for (size_t b = 0; b < grids.size(); ++b) {
        const Arr& uo = uold[b];
        const Arr& un = unew[b];
        const Arr& um = umac[0][b];
        const Arr& vm = umac[1][b];
        const Arr& ex = uedge[0][b];
        const Arr& ey = uedge[1][b];
        const Arr& fo = force[b];
        Arr w0, wm, ez, sp;
        if (!c.p.spherical) w0 = w0_cart[b];
        if (dim == 3) {
            wm = umac[2][b];
            ez = uedge[2][b];
        }
        if (do_sponge) sp = sponge[b];
        if (!c.p.spherical) {
            par_for(grids[b], [&](int i, int j, int k) {
                const double ubar = 0.5 * (um(i, j, k) + um(i + 1, j, k));
                const double vbar = 0.5 * (vm(i, j, k) + vm(i, j + 1, k));
                double g[3];
                for (int n = 0; n < dim; ++n) {
                    g[n] = ubar * (ex(i + 1, j, k, n) - ex(i, j, k, n)) / dx[0] +
                           vbar * (ey(i, j + 1, k, n) - ey(i, j, k, n)) / dx[1];
                    if (dim == 3) {
                        const double wbar = 0.5 * (wm(i, j, k) + wm(i, j, k + 1));
                        g[n] = g[n] + wbar * (ez(i, j, k + 1, n) - ez(i, j, k, n)) / dx[2];
                    }
                }
                for (int n = 0; n < dim; ++n)
                    un(i, j, k, n) = uo(i, j, k, n) - dt_loc * g[n] + dt_loc * fo(i, j, k, n);
                // (w0 . grad) Utilde with w0 the cell-centred w0_cart of this
                // cell and the one above it
                double w0bar;
                if (dim == 2)
                    w0bar = 0.5 * (w0(i, j, k, dim - 1) + w0(i, j + 1, k, dim - 1));
                else
                    w0bar = 0.5 * (w0(i, j, k, dim - 1) + w0(i, j, k + 1, dim - 1));
                for (int n = 0; n < dim; ++n) {
                    if (dim == 2)
                        un(i, j, k, n) -= dt_loc * w0bar * (ey(i, j + 1, k, n) - ey(i, j, k, n)) / dx[1];
                    else
                        un(i, j, k, n) -= dt_loc * w0bar * (ez(i, j, k + 1, n) - ez(i, j, k, n)) / dx[2];
                    if (do_sponge) un(i, j, k, n) *= sp(i, j, k);
                }
            });
        } else {
            const Arr& w0x = w0mac[0][b];
            const Arr& w0y = w0mac[1][b];
            const Arr& w0z = w0mac[2][b];
            par_for(grids[b], [&](int i, int j, int k) {
                const double ubar = 0.5 * (um(i, j, k) + um(i + 1, j, k));
                const double vbar = 0.5 * (vm(i, j, k) + vm(i, j + 1, k));
                const double wbar = 0.5 * (wm(i, j, k) + wm(i, j, k + 1));
                double g[3], gx[3], gy[3], gz[3];
                for (int n = 0; n < 3; ++n) {
                    g[n] = ubar * (ex(i + 1, j, k, n) - ex(i, j, k, n)) / dx[0] +
                           vbar * (ey(i, j + 1, k, n) - ey(i, j, k, n)) / dx[1] +
                           wbar * (ez(i, j, k + 1, n) - ez(i, j, k, n)) / dx[2];
                }
                for (int n = 0; n < 3; ++n)
                    un(i, j, k, n) = uo(i, j, k, n) - dt_loc * g[n] + dt_loc * fo(i, j, k, n);
                for (int n = 0; n < 3; ++n) {
                    gx[n] = (ex(i + 1, j, k, n) - ex(i, j, k, n)) / dx[0];
                    gy[n] = (ey(i, j + 1, k, n) - ey(i, j, k, n)) / dx[1];
                    gz[n] = (ez(i, j, k + 1, n) - ez(i, j, k, n)) / dx[2];
                }
                for (int n = 0; n < 3; ++n) {
                    const double w0g = gx[n] * 0.5 * (w0x(i, j, k) + w0x(i + 1, j, k)) +
                                       gy[n] * 0.5 * (w0y(i, j, k) + w0y(i, j + 1, k)) +
                                       gz[n] * 0.5 * (w0z(i, j, k) + w0z(i, j, k + 1));
                    un(i, j, k, n) -= dt_loc * w0g;
                }
                if (do_sponge)
                    for (int n = 0; n < 3; ++n) un(i, j, k, n) *= sp(i, j, k);
            });
        }
    }
    return MX_OK;
}

}  // namespace mxo
