// oracle_ppm.cpp -- CPU oracle: Maestro::PPM and Maestro::Slope{x,y,z}.
//
// TEST INFRASTRUCTURE ONLY (see oracle_core.H).
//
// Restates Source/MaestroPPM.cpp:6-2263 and Source/MaestroSlopes.cpp:6-600.
// The reference writes one ParallelFor body per direction; here one body is
// written once and parameterised by the direction, with the places where the
// reference's x, y and z bodies differ kept as explicit per-direction flags:
//   - ppm_type 2, lo-side boundary re-limit, bigm branch: x tests
//     `>= 1.e10` (MaestroPPM.cpp:566), y and z `>= 1.e-10` (:1310, :2054);
//   - ppm_type 2, hi side at domhi-1: y and z first set sedgerr = s(+2)
//     (:1342, :2086), x does not (:598);
//   - ppm_type 2, lo side at domlo: y sets sedge = s(-1) (:1192);
//   - slope_order 4 boundary cells: Slopex selects `dlim` where Slopey and
//     Slopez select `slim` (MaestroSlopes.cpp:142,156,178,192 vs :341,...).
// Arithmetic expressions keep the reference's operand order so that results
// are reproducible bit for bit.
#include <cmath>

#include "oracle_core.H"

namespace mxo {

namespace {

// Colella & Woodward quadratic limiter (MaestroPPM.cpp:106-115)
inline void cw_limit(double s0, double& sm, double& sp) {
    if ((sp - s0) * (s0 - sm) <= 0.0) {
        sp = s0;
        sm = s0;
    } else if (ab(sp - s0) >= 2.0 * ab(sm - s0)) {
        sp = 3.0 * s0 - 2.0 * sm;
    } else if (ab(sm - s0) >= 2.0 * ab(sp - s0)) {
        sm = 3.0 * s0 - 2.0 * sp;
    }
}

// van Leer slope used by ppm_type 1 (MaestroPPM.cpp:40-57)
inline double vl_slope(double a, double b, double c) {
    // a = s(m-1), b = s(m), c = s(m+1)
    double dsc = 0.5 * (c - a);
    double dsl = 2.0 * (b - a);
    double dsr = 2.0 * (c - b);
    double r = 0.0;
    if (dsl * dsr > 0.0) r = cs(1.0, dsc) * mn(ab(dsc), mn(ab(dsl), ab(dsr)));
    return r;
}

// 4th-order face value with the Colella-Sekora limiter on the face
// (MaestroPPM.cpp:256-350); face between sb and sc, sa and sd outside.
inline double cs_face(double sa, double sb, double sc, double sd) {
    const double C = 1.25;
    double e = (7.0 / 12.0) * (sb + sc) - (1.0 / 12.0) * (sa + sd);
    if ((e - sb) * (sc - e) < 0.0) {
        double D2 = 3.0 * (sb - 2.0 * e + sc);
        double D2L = sa - 2.0 * sb + sc;
        double D2R = sb - 2.0 * sc + sd;
        double sgn = cs(1.0, D2);
        double D2LIM =
            sgn * mx(mn(C * sgn * D2L, mn(C * sgn * D2R, sgn * D2)), 0.0);
        e = 0.5 * (sb + sc) - D2LIM / 6.0;
    }
    return e;
}

// Colella 2008 limiter on the cell (MaestroPPM.cpp:352-441); S(o) = s(m+o).
template <class SF>
inline void cs_cell(const SF& S, double sedgel, double sedge, double sedger,
                    double sedgerr, double bigm_thresh, double& sm,
                    double& sp) {
    const double C = 1.25;
    double alphap = sedger - S(0);
    double alpham = sedge - S(0);
    bool bigp = ab(alphap) > 2.0 * ab(alpham);
    bool bigm = ab(alpham) > 2.0 * ab(alphap);
    bool extremum = false;

    if (alpham * alphap >= 0.0) {
        extremum = true;
    } else if (bigp || bigm) {
        double dafacem = sedge - sedgel;
        double dafacep = sedgerr - sedger;
        double dabarm = S(0) - S(-1);
        double dabarp = S(1) - S(0);
        double dafacemin = mn(ab(dafacem), ab(dafacep));
        double dabarmin = mn(ab(dabarm), ab(dabarp));
        double dachkm = 0.0;
        double dachkp = 0.0;
        if (dafacemin >= dabarmin) {
            dachkm = dafacem;
            dachkp = dafacep;
        } else {
            dachkm = dabarm;
            dachkp = dabarp;
        }
        extremum = (dachkm * dachkp <= 0.0);
    }

    if (extremum) {
        double D2 = 6.0 * (alpham + alphap);
        double D2L = S(-2) - 2.0 * S(-1) + S(0);
        double D2R = S(0) - 2.0 * S(1) + S(2);
        double D2C = S(-1) - 2.0 * S(0) + S(1);
        double sgn = cs(1.0, D2);
        double D2LIM = mx(
            mn(sgn * D2, mn(C * sgn * D2L, mn(C * sgn * D2R, C * sgn * D2C))),
            0.0);
        double D2ABS = mx(ab(D2), 1.e-10);
        alpham = alpham * D2LIM / D2ABS;
        alphap = alphap * D2LIM / D2ABS;
    } else {
        if (bigp) {
            double sgn = cs(1.0, alpham);
            double amax = -alphap * alphap / (4.0 * (alpham + alphap));
            double delam = S(-1) - S(0);
            if (sgn * amax >= sgn * delam) {
                if (sgn * (delam - alpham) >= 1.e-10) {
                    alphap = -2.0 * delam -
                             2.0 * sgn * std::sqrt(delam * delam - delam * alpham);
                } else {
                    alphap = -2.0 * alpham;
                }
            }
        }
        if (bigm) {
            double sgn = cs(1.0, alphap);
            double amax = -alpham * alpham / (4.0 * (alpham + alphap));
            double delap = S(1) - S(0);
            if (sgn * amax >= sgn * delap) {
                if (sgn * (delap - alphap) >= bigm_thresh) {
                    alpham = -2.0 * delap -
                             2.0 * sgn * std::sqrt(delap * delap - delap * alphap);
                } else {
                    alpham = -2.0 * alphap;
                }
            }
        }
    }
    sm = S(0) + alpham;
    sp = S(0) + alphap;
}

// boundary face value "-0.2 a + 0.75 b + 0.5 c - 0.05 d" (MaestroPPM.cpp:124)
inline double bnd4(double a, double b, double c, double d) {
    return -0.2 * a + 0.75 * b + 0.5 * c - 0.05 * d;
}

}  // namespace

// One direction of Maestro::PPM for one cell: edge values sm, sp of s along
// direction `d` at position `pos` (index along d).
template <class SF>
static void ppm_edges(int ppm_type, int d, const SF& S, int pos, int dlo,
                      int dhi, int bclo, int bchi, double& sm_out,
                      double& sp_out) {
    double sm = 0.0, sp = 0.0;
    if (ppm_type == 1) {
        // MaestroPPM.cpp:34-205 (x), :780-951 (y), :1523-1694 (z)
        double dsvl_l = vl_slope(S(-2), S(-1), S(0));
        double dsvl_r = vl_slope(S(-1), S(0), S(1));
        sm = 0.5 * (S(0) + S(-1)) - (dsvl_r - dsvl_l) / 6.0;
        sm = mx(sm, mn(S(0), S(-1)));
        sm = mn(sm, mx(S(0), S(-1)));

        dsvl_l = vl_slope(S(-1), S(0), S(1));
        dsvl_r = vl_slope(S(0), S(1), S(2));
        sp = 0.5 * (S(1) + S(0)) - (dsvl_r - dsvl_l) / 6.0;
        sp = mx(sp, mn(S(1), S(0)));
        sp = mn(sp, mx(S(1), S(0)));

        double sedgel = sp;
        double sedger = sm;

        cw_limit(S(0), sm, sp);

        const bool lo_ho = (bclo == MX_BC_EXT_DIR || bclo == MX_BC_HOEXTRAP);
        const bool hi_ho = (bchi == MX_BC_EXT_DIR || bchi == MX_BC_HOEXTRAP);
        if (pos == dlo) {
            if (lo_ho) {
                sm = S(-1);
                sp = bnd4(S(-1), S(0), S(1), S(2));
                sp = mx(sp, mn(S(1), S(0)));
                sp = mn(sp, mx(S(1), S(0)));
            }
        } else if (pos == dlo + 1) {
            if (lo_ho) {
                sm = bnd4(S(-2), S(-1), S(0), S(1));
                sm = mx(sm, mn(S(0), S(-1)));
                sm = mn(sm, mx(S(0), S(-1)));
                sp = sedgel;
                cw_limit(S(0), sm, sp);
            }
        } else if (pos == dhi) {
            if (hi_ho) {
                sp = S(1);
                sm = bnd4(S(1), S(0), S(-1), S(-2));
                sm = mx(sm, mn(S(-1), S(0)));
                sm = mn(sm, mx(S(-1), S(0)));
            }
        } else if (pos == dhi - 1) {
            if (hi_ho) {
                sp = bnd4(S(2), S(1), S(0), S(-1));
                sp = mx(sp, mn(S(0), S(1)));
                sp = mn(sp, mx(S(0), S(1)));
                sm = sedger;
                cw_limit(S(0), sm, sp);
            }
        }
    } else {
        // ppm_type 2: MaestroPPM.cpp:255-719 (x), :1001-1463 (y), :1747-2207 (z)
        double sedgel = cs_face(S(-3), S(-2), S(-1), S(0));
        double sedge = cs_face(S(-2), S(-1), S(0), S(1));
        double sedger = cs_face(S(-1), S(0), S(1), S(2));
        double sedgerr = cs_face(S(0), S(1), S(2), S(3));

        cs_cell(S, sedgel, sedge, sedger, sedgerr, 1.e-10, sm, sp);

        if (bclo == MX_BC_EXT_DIR || bclo == MX_BC_HOEXTRAP) {
            if (pos == dlo) {
                if (d == 1) sedge = S(-1);  // y only; no effect on sm/sp
                sm = S(-1);
                sp = bnd4(S(-1), S(0), S(1), S(2));
                sp = mx(sp, mn(S(1), S(0)));
                sp = mn(sp, mx(S(1), S(0)));
            } else if (pos == dlo + 1) {
                sedgel = S(-2);
                sedge = bnd4(S(-2), S(-1), S(0), S(1));
                sedge = mx(sedge, mn(S(0), S(-1)));
                sedge = mn(sedge, mx(S(0), S(-1)));
            } else if (pos == dlo + 2) {
                sedgel = bnd4(S(-3), S(-2), S(-1), S(0));
                sedgel = mx(sedgel, mn(S(-1), S(-2)));
                sedgel = mn(sedgel, mx(S(-1), S(-2)));
            }
            if (pos == dlo + 1 || pos == dlo + 2) {
                const double thr = (d == 0) ? 1.e10 : 1.e-10;
                cs_cell(S, sedgel, sedge, sedger, sedgerr, thr, sm, sp);
            }
        }

        if (bchi == MX_BC_EXT_DIR || bchi == MX_BC_HOEXTRAP) {
            if (pos == dhi) {
                sp = S(1);
                sm = bnd4(S(1), S(0), S(-1), S(-2));
                sm = mx(sm, mn(S(-1), S(0)));
                sm = mn(sm, mx(S(-1), S(0)));
            } else if (pos == dhi - 1) {
                if (d != 0) sedgerr = S(2);  // y and z only
                sedger = bnd4(S(2), S(1), S(0), S(-1));
                sedger = mx(sedger, mn(S(0), S(1)));
                sedger = mn(sedger, mx(S(0), S(1)));
            } else if (pos == dhi - 2) {
                sedgerr = bnd4(S(3), S(2), S(1), S(0));
                sedgerr = mx(sedgerr, mn(S(1), S(2)));
                sedgerr = mn(sedgerr, mx(S(1), S(2)));
            }
            if (pos == dhi - 1 || pos == dhi - 2) {
                cs_cell(S, sedgel, sedge, sedger, sedgerr, 1.e-10, sm, sp);
            }
        }
    }
    sm_out = sm;
    sp_out = sp;
}

// Maestro::PPM (MaestroPPM.cpp:6-2263)
void PPM(const Ctx& c, const IBox& bx, const Arr& s, const Arr& u,
         const Arr& v, const Arr& w, const Arr& Ip, const Arr& Im,
         const IBox& dom, const mx_bcrec* bcs, const double* dx, bool is_umac,
         int comp, int bccomp) {
    const int n = comp;
    const double dt_local = c.p.dt;
    const double rel_eps_local = c.p.rel_eps;
    const int ppm_type = c.p.ppm_type;

    for (int d = 0; d < c.dim(); ++d) {
        const int bclo = bcs[bccomp].lo[d];
        const int bchi = bcs[bccomp].hi[d];
        const Arr& vel = (d == 0) ? u : (d == 1) ? v : w;
        const int e0 = (d == 0), e1 = (d == 1), e2 = (d == 2);
        const int dlo = dom.lo[d], dhi = dom.hi[d];
        const double dxd = dx[d];

        par_for(bx, [&](int i, int j, int k) {
            auto S = [&](int o) { return s(i + o * e0, j + o * e1, k + o * e2, n); };
            const int pos = (d == 0) ? i : (d == 1) ? j : k;
            double sm, sp;
            ppm_edges(ppm_type, d, S, pos, dlo, dhi, bclo, bchi, sm, sp);

            const double s0 = S(0);
            double s6 = 6.0 * s0 - 3.0 * (sm + sp);

            if (is_umac) {
                const double uhi = vel(i + e0, j + e1, k + e2);
                const double ulo = vel(i, j, k);
                double sigma = ab(uhi) * dt_local / dxd;
                if (uhi > rel_eps_local) {
                    Ip(i, j, k, d) =
                        sp - 0.5 * sigma * (sp - sm - (1.0 - 2.0 / 3.0 * sigma) * s6);
                } else {
                    Ip(i, j, k, d) = s0;
                }
                sigma = ab(ulo) * dt_local / dxd;
                if (ulo < -rel_eps_local) {
                    Im(i, j, k, d) =
                        sm + 0.5 * sigma * (sp - sm + (1.0 - 2.0 / 3.0 * sigma) * s6);
                } else {
                    Im(i, j, k, d) = s0;
                }
            } else {
                const double uc = vel(i, j, k);
                double sigma = ab(uc) * dt_local / dxd;
                if (uc > rel_eps_local) {
                    Ip(i, j, k, d) =
                        sp - 0.5 * sigma * (sp - sm - (1.0 - 2.0 / 3.0 * sigma) * s6);
                } else {
                    Ip(i, j, k, d) = s0;
                }
                sigma = ab(uc) * dt_local / dxd;
                if (uc < -rel_eps_local) {
                    Im(i, j, k, d) =
                        sm + 0.5 * sigma * (sp - sm + (1.0 - 2.0 / 3.0 * sigma) * s6);
                } else {
                    Im(i, j, k, d) = s0;
                }
            }
        });
    }
}

// Maestro::Slopex/Slopey/Slopez (MaestroSlopes.cpp:6-600)
void Slope(int dir, int slope_order, const IBox& bx, const Arr& s,
           const Arr& sl, const IBox& dom, const mx_bcrec* bcs, int ncomp,
           int bc_start_comp) {
    const int ilo = dom.lo[dir];
    const int ihi = dom.hi[dir];
    const int e0 = (dir == 0), e1 = (dir == 1), e2 = (dir == 2);

    for (int n = 0; n < ncomp; ++n) {
        const int bclo = bcs[bc_start_comp + n].lo[dir];
        const int bchi = bcs[bc_start_comp + n].hi[dir];
        const bool lo_ho = (bclo == MX_BC_EXT_DIR || bclo == MX_BC_HOEXTRAP);
        const bool hi_ho = (bchi == MX_BC_EXT_DIR || bchi == MX_BC_HOEXTRAP);

        par_for(bx, [&](int i, int j, int k) {
            auto S = [&](int o) { return s(i + o * e0, j + o * e1, k + o * e2, n); };
            const int pos = (dir == 0) ? i : (dir == 1) ? j : k;
            double& out = sl(i, j, k, n);

            if (slope_order == 0) {
                out = 0.0;
            } else if (slope_order == 2) {
                double del = 0.5 * (S(1) - S(-1));
                double dpls = 2.0 * (S(1) - S(0));
                double dmin = 2.0 * (S(0) - S(-1));
                double slim = mn(ab(dpls), ab(dmin));
                slim = dpls * dmin > 0.0 ? slim : 0.0;
                double sflag = cs(1.0, del);
                out = sflag * mn(slim, ab(del));

                if (lo_ho) {
                    if (pos == ilo - 1) {
                        out = 0.0;
                    } else if (pos == ilo) {
                        del = (S(1) + 3.0 * S(0) - 4.0 * S(-1)) / 3.0;
                        dpls = 2.0 * (S(1) - S(0));
                        dmin = 2.0 * (S(0) - S(-1));
                        slim = mn(ab(dpls), ab(dmin));
                        slim = dpls * dmin > 0.0 ? slim : 0.0;
                        sflag = cs(1.0, del);
                        out = sflag * mn(slim, ab(del));
                    }
                }
                if (hi_ho) {
                    if (pos == ihi + 1) {
                        out = 0.0;
                    } else if (pos == ihi) {
                        del = -(S(-1) + 3.0 * S(0) - 4.0 * S(1)) / 3.0;
                        if (dir == 0) {
                            dpls = 2.0 * (S(0) - S(-1));
                            dmin = 2.0 * (S(1) - S(0));
                        } else {
                            dpls = 2.0 * (S(1) - S(0));
                            dmin = 2.0 * (S(0) - S(-1));
                        }
                        slim = mn(ab(dpls), ab(dmin));
                        slim = dpls * dmin > 0.0 ? slim : 0.0;
                        sflag = cs(1.0, del);
                        out = sflag * mn(slim, ab(del));
                    }
                }
            } else if (slope_order == 4) {
                // left
                double dcen = 0.5 * (S(0) - S(-2));
                double dmin = 2.0 * (S(-1) - S(-2));
                double dpls = 2.0 * (S(0) - S(-1));
                double dlim = mn(ab(dmin), ab(dpls));
                dlim = dpls * dmin > 0.0 ? dlim : 0.0;
                double dflag = cs(1.0, dcen);
                double dxl = dflag * mn(dlim, ab(dcen));
                // right
                dcen = 0.5 * (S(2) - S(0));
                dmin = 2.0 * (S(1) - S(0));
                dpls = 2.0 * (S(2) - S(1));
                dlim = mn(ab(dmin), ab(dpls));
                dlim = dpls * dmin > 0.0 ? dlim : 0.0;
                dflag = cs(1.0, dcen);
                double dxr = dflag * mn(dlim, ab(dcen));
                // center
                dcen = 0.5 * (S(1) - S(-1));
                dmin = 2.0 * (S(0) - S(-1));
                dpls = 2.0 * (S(1) - S(0));
                dlim = mn(ab(dmin), ab(dpls));
                dlim = dpls * dmin > 0.0 ? dlim : 0.0;
                dflag = cs(1.0, dcen);

                double ds = 4.0 / 3.0 * dcen - (dxr + dxl) / 6.0;
                out = dflag * mn(ab(ds), dlim);

                // x keeps the centre dlim in the boundary limiters, y/z slim
                auto pick = [&](double slim_v) { return dir == 0 ? dlim : slim_v; };

                if (lo_ho) {
                    if (pos == ilo - 1) {
                        out = 0.0;
                    } else if (pos == ilo) {
                        double del = -16.0 / 15.0 * S(-1) + 0.5 * S(0) +
                                     2.0 / 3.0 * S(1) - 0.1 * S(2);
                        dmin = 2.0 * (S(0) - S(-1));
                        dpls = 2.0 * (S(1) - S(0));
                        double slim = mn(ab(dpls), ab(dmin));
                        slim = dpls * dmin > 0.0 ? pick(slim) : 0.0;
                        double sflag = cs(1.0, del);
                        out = sflag * mn(slim, ab(del));
                    } else if (pos == ilo + 1) {
                        double del = -16.0 / 15.0 * S(-2) + 0.5 * S(-1) +
                                     2.0 / 3.0 * S(0) - 0.1 * S(1);
                        dmin = 2.0 * (S(-1) - S(-2));
                        dpls = 2.0 * (S(0) - S(-1));
                        double slim = mn(ab(dpls), ab(dmin));
                        slim = dpls * dmin > 0.0 ? pick(slim) : 0.0;
                        double sflag = cs(1.0, del);
                        dxl = sflag * mn(slim, ab(del));
                        ds = 4.0 / 3.0 * dcen - (dxr + dxl) / 6.0;
                        out = dflag * mn(ab(ds), dlim);
                    }
                }
                if (hi_ho) {
                    if (pos == ihi + 1) {
                        out = 0.0;
                    } else if (pos == ihi) {
                        double del = -(-16.0 / 15.0 * S(1) + 0.5 * S(0) +
                                       2.0 / 3.0 * S(-1) - 0.1 * S(-2));
                        dmin = 2.0 * (S(0) - S(-1));
                        dpls = 2.0 * (S(1) - S(0));
                        double slim = mn(ab(dpls), ab(dmin));
                        slim = dpls * dmin > 0.0 ? pick(slim) : 0.0;
                        double sflag = cs(1.0, del);
                        out = sflag * mn(slim, ab(del));
                    } else if (pos == ihi - 1) {
                        double del = -(-16.0 / 15.0 * S(2) + 0.5 * S(1) +
                                       2.0 / 3.0 * S(0) - 0.1 * S(-1));
                        dmin = 2.0 * (S(1) - S(0));
                        dpls = 2.0 * (S(2) - S(1));
                        double slim = mn(ab(dpls), ab(dmin));
                        slim = dpls * dmin > 0.0 ? pick(slim) : 0.0;
                        double sflag = cs(1.0, del);
                        dxr = sflag * mn(slim, ab(del));
                        ds = 4.0 / 3.0 * dcen - (dxl + dxr) / 6.0;
                        out = dflag * mn(ab(ds), dlim);
                    }
                }
            }
        });
    }
}

}  // namespace mxo
