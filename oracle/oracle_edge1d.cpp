// oracle_edge1d.cpp -- CPU oracle of Maestro::MakeEdgeState1d.
//
// TEST INFRASTRUCTURE ONLY (see oracle_core.H).  A line-by-line restatement
// of MaestroMakeEdgeState.cpp: the spherical variant (:17-395) on radial
// level 0 with three reflected / extrapolated ghost cells, and the planar
// variant (:397-981) over radial levels and disjoint chunks with the
// coarse-fine edge synchronisation.  Temporaries start at zero, as the
// reference's BaseState constructor fills them (BaseState.H:308-316), so
// reads of entries a chunk does not compute see the same values.
#include <cmath>
#include <vector>

#include "oracle_core.H"

namespace mxo {

namespace {

inline double mn3(double a, double b, double c) { return mn(mn(a, b), c); }

struct Row {
    const double* p;
    long long len;
    double operator()(int n, int r) const { return p[(long long)n * len + r]; }
};
struct RowW {
    double* p;
    long long len;
    double& operator()(int n, int r) const { return p[(long long)n * len + r]; }
};

// van Leer slope of the spherical ghosted array at g (:126-138)
inline double vl_g(const std::vector<double>& sg, int g) {
    double del = 0.5 * (sg[g + 1] - sg[g - 1]);
    double dmin = 2.0 * (sg[g] - sg[g - 1]);
    double dpls = 2.0 * (sg[g + 1] - sg[g]);
    return dmin * dpls > 0.0 ? cs(1.0, del) * mn3(ab(del), ab(dmin), ab(dpls)) : 0.0;
}

// Colella-Sekora face value between g-1 and g (:224-246)
inline double cs_face_g(const std::vector<double>& sg, int g) {
    const double C = 1.25;
    double e = (7.0 / 12.0) * (sg[g - 1] + sg[g]) - (1.0 / 12.0) * (sg[g - 2] + sg[g + 1]);
    if ((e - sg[g - 1]) * (sg[g] - e) < 0.0) {
        double D2 = 3.0 * (sg[g - 1] - 2.0 * e + sg[g]);
        double D2L = sg[g - 2] - 2.0 * sg[g - 1] + sg[g];
        double D2R = sg[g - 1] - 2.0 * sg[g] + sg[g + 1];
        double sgn = cs(1.0, D2);
        double D2LIM = sgn * mx(mn(C * sgn * D2L, mn(C * sgn * D2R, sgn * D2)), 0.0);
        e = 0.5 * (sg[g - 1] + sg[g]) - D2LIM / 6.0;
    }
    return e;
}

// Colella 2008 cell limiter (:250-339 / :785-876); sm1, s0, sp1 = s(r-1),
// s(r), s(r+1), sm2, sp2 = s(r-2), s(r+2); fl..frr = faces r-1 .. r+2
inline void cs08(double sm2, double sm1, double s0, double sp1, double sp2, double fl, double f,
                 double fr, double frr, double& sm, double& sp) {
    const double C = 1.25;
    double alphap = fr - s0;
    double alpham = f - s0;
    bool bigp = ab(alphap) > 2.0 * ab(alpham);
    bool bigm = ab(alpham) > 2.0 * ab(alphap);
    bool extremum = false;
    if (alpham * alphap >= 0.0) {
        extremum = true;
    } else if (bigp || bigm) {
        double dafacem = f - fl;
        double dafacep = frr - fr;
        double dabarm = s0 - sm1;
        double dabarp = sp1 - s0;
        double dafacemin = mn(ab(dafacem), ab(dafacep));
        double dabarmin = mn(ab(dabarm), ab(dabarp));
        double dachkm = 0.0, dachkp = 0.0;
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
        double D2L = sm2 - 2.0 * sm1 + s0;
        double D2R = s0 - 2.0 * sp1 + sp2;
        double D2C = sm1 - 2.0 * s0 + sp1;
        double sgn = cs(1.0, D2);
        double D2LIM =
            mx(mn(sgn * D2, mn(C * sgn * D2L, mn(C * sgn * D2R, C * sgn * D2C))), 0.0);
        double D2ABS = mx(ab(D2), 1.e-10);
        alpham = alpham * D2LIM / D2ABS;
        alphap = alphap * D2LIM / D2ABS;
    } else {
        if (bigp) {
            double sgn = cs(1.0, alpham);
            double amax = -alphap * alphap / (4.0 * (alpham + alphap));
            double delam = sm1 - s0;
            if (sgn * amax >= sgn * delam) {
                if (sgn * (delam - alpham) >= 1.e-10) {
                    alphap = (-2.0 * delam - 2.0 * sgn * std::sqrt(delam * delam - delam * alpham));
                } else {
                    alphap = -2.0 * alpham;
                }
            }
        }
        if (bigm) {
            double sgn = cs(1.0, alphap);
            double amax = -alpham * alpham / (4.0 * (alpham + alphap));
            double delap = sp1 - s0;
            if (sgn * amax >= sgn * delap) {
                if (sgn * (delap - alphap) >= 1.e-10) {
                    alpham = (-2.0 * delap - 2.0 * sgn * std::sqrt(delap * delap - delap * alphap));
                } else {
                    alpham = -2.0 * alphap;
                }
            }
        }
    }
    sm = s0 + alpham;
    sp = s0 + alphap;
}

// quadratic (CW84) limiter (:187-197)
inline void cw84(double s0, double& sm, double& sp) {
    if ((sp - s0) * (s0 - sm) <= 0.0) {
        sp = s0;
        sm = s0;
    } else if (ab(sp - s0) >= 2.0 * ab(sm - s0)) {
        sp = 3.0 * s0 - 2.0 * sm;
    } else if (ab(sm - s0) >= 2.0 * ab(sp - s0)) {
        sm = 3.0 * s0 - 2.0 * sp;
    }
}

// one-sided / limited van Leer slope of planar level data at rr (:597-614)
inline double vl_planar(const Row& s, int n, int rr, int nr) {
    double d = 0.0;
    if (rr == 0) {
        d = s(n, rr + 1) - s(n, rr);
    } else if (rr == nr - 1) {
        d = s(n, rr) - s(n, rr - 1);
    } else if (rr > 0 && rr < nr - 1) {
        double del = 0.5 * (s(n, rr + 1) - s(n, rr - 1));
        double dmin = 2.0 * (s(n, rr) - s(n, rr - 1));
        double dpls = 2.0 * (s(n, rr + 1) - s(n, rr));
        d = dmin * dpls > 0.0 ? cs(1.0, del) * mn(ab(del), mn(ab(dmin), ab(dpls))) : 0.0;
    }
    return d;
}

// limited Fromm slope of planar level data at rr (:479-495)
inline double fromm_planar(const Row& s, int n, int rr, int nr) {
    double d = 0.0;
    if (rr == 0) {
        d = s(n, rr + 1) - s(n, rr);
    } else if (rr == nr - 1) {
        d = s(n, rr) - s(n, rr - 1);
    } else if (rr > 0 && rr < nr - 1) {
        double cen = 0.5 * (s(n, rr + 1) - s(n, rr - 1));
        double dpls = 2.0 * (s(n, rr + 1) - s(n, rr));
        double dmin = 2.0 * (s(n, rr) - s(n, rr - 1));
        double lim = mn(ab(dmin), ab(dpls));
        lim = dpls * dmin > 0.0 ? lim : 0.0;
        double flag = cs(1.0, cen);
        d = flag * mn(lim, ab(cen));
    }
    return d;
}

void edge1d_sphr(const Ctx* c, const mx_radial_geom* rg, const mx_basestate* s_bs,
                 mx_basestate* sedge_bs, const mx_basestate* f_bs, const mx_basestate* w0_bs) {
    const mx_adv_params& p = c->p;
    const double rel_eps = p.rel_eps;
    const double dt = p.dt;
    const double dth = 0.5 * dt;
    const double dr0 = rg->dr[0];
    const double dtdr = dt / dr0;
    const int nr = rg->nr_fine;
    Row s{s_bs->p, s_bs->len}, force{f_bs->p, f_bs->len}, w0{w0_bs->p, w0_bs->len};
    RowW sedge{sedge_bs->p, sedge_bs->len};
    std::vector<double> sedgel(nr + 1, 0.0), sedger(nr + 1, 0.0);
    const int ng = 3;
    std::vector<double> sg(nr + 2 * ng, 0.0);
    for (int i = 0; i < nr; ++i) sg[i + ng] = s(0, i);
    for (int i = 0; i < ng; ++i) {
        sg[ng - 1 - i] = s(0, i);
        sg[ng + nr + i] = s(0, nr - 1);
    }
    for (int r = 0; r < nr; ++r) {
        if (p.ppm_type == 0) {
            double slope = 0.0;
            if (p.slope_order == 0) {
                slope = 0.0;
            } else if (p.slope_order == 2) {
                int g = r + ng;
                double del = 0.5 * (sg[g + 1] - sg[g - 1]);
                double dpls = 2.0 * (sg[g + 1] - sg[g]);
                double dmin = 2.0 * (sg[g] - sg[g - 1]);
                double slim = mn(ab(dpls), ab(dmin));
                slim = dpls * dmin > 0.0 ? slim : 0.0;
                double sflag = cs(1.0, del);
                slope = sflag * mn(slim, ab(del));
            } else if (p.slope_order == 4) {
                double cen[3], lim[3], flag[3], fromm[3];
                for (int i = 0; i < 3; ++i) {
                    int g = r + ng - 1 + i;
                    cen[i] = 0.5 * (sg[g + 1] - sg[g - 1]);
                    double dpls = 2.0 * (sg[g + 1] - sg[g]);
                    double dmin = 2.0 * (sg[g] - sg[g - 1]);
                    lim[i] = mn(ab(dmin), ab(dpls));
                    lim[i] = dpls * dmin > 0.0 ? lim[i] : 0.0;
                    flag[i] = cs(1.0, cen[i]);
                    fromm[i] = flag[i] * mn(lim[i], ab(cen[i]));
                }
                double ds = 4.0 / 3.0 * cen[1] - (fromm[2] + fromm[0]) / 6.0;
                slope = flag[1] * mn(ab(ds), lim[1]);
            }
            double u = 0.5 * (w0(0, r) + w0(0, r + 1));
            double ubardth = dth * u / dr0;
            sedgel[r + 1] = s(0, r) + (0.5 - ubardth) * slope + dth * force(0, r);
            sedger[r] = s(0, r) - (0.5 + ubardth) * slope + dth * force(0, r);
            continue;
        }
        double sm, sp;
        const int g = r + ng;
        if (p.ppm_type == 1) {
            double dsscrr = vl_g(sg, g);
            double dsscrl = vl_g(sg, g - 1);
            sm = 0.5 * (sg[g] + sg[g - 1]) - (dsscrr - dsscrl) / 6.0;
            sm = mx(sm, mn(sg[g], sg[g - 1]));
            sm = mn(sm, mx(sg[g], sg[g - 1]));
            dsscrl = dsscrr;
            dsscrr = vl_g(sg, g + 1);
            sp = 0.5 * (sg[g + 1] + sg[g]) - (dsscrr - dsscrl) / 6.0;
            sp = mx(sp, mn(sg[g + 1], sg[g]));
            sp = mn(sp, mx(sg[g + 1], sg[g]));
            cw84(s(0, r), sm, sp);
        } else {
            double sedget[4];
            for (int j = 0; j < 4; ++j) sedget[j] = cs_face_g(sg, r + ng - 1 + j);
            cs08(sg[g - 2], sg[g - 1], sg[g], sg[g + 1], sg[g + 2], sedget[0], sedget[1],
                 sedget[2], sedget[3], sm, sp);
        }
        double sigmap = ab(w0(0, r + 1)) * dtdr;
        double sigmam = ab(w0(0, r)) * dtdr;
        double s6 = 6.0 * s(0, r) - 3.0 * (sm + sp);
        double Ip = w0(0, r + 1) > rel_eps
                        ? sp - 0.5 * sigmap * (sp - sm - (1.0 - 2.0 / 3.0 * sigmap) * s6)
                        : s(0, r);
        double Im = w0(0, r) < -rel_eps
                        ? sm + 0.5 * sigmam * (sp - sm + (1.0 - 2.0 / 3.0 * sigmam) * s6)
                        : s(0, r);
        sedgel[r + 1] = Ip + dth * force(0, r);
        sedger[r] = Im + dth * force(0, r);
    }
    for (int r = 0; r <= nr; ++r) {
        if (r == 0) {
            sedgel[r] = sedger[r];
        } else if (r == nr) {
            sedger[r] = sedgel[r];
        }
        sedge(0, r) = w0(0, r) > 0.0 ? sedgel[r] : sedger[r];
        sedge(0, r) = ab(w0(0, r)) < rel_eps ? 0.5 * (sedger[r] + sedgel[r]) : sedge(0, r);
    }
}

void edge1d_planar(const Ctx* c, const mx_radial_geom* rg, const mx_basestate* s_bs,
                   mx_basestate* sedge_bs, const mx_basestate* f_bs,
                   const mx_basestate* w0_bs) {
    const mx_adv_params& p = c->p;
    const double rel_eps = p.rel_eps;
    const double dt = p.dt;
    const double dth = 0.5 * dt;
    const int nlev = rg->nlev;
    const long long L = rg->nr_fine + 1;
    Row s{s_bs->p, s_bs->len}, force{f_bs->p, f_bs->len}, w0{w0_bs->p, w0_bs->len};
    RowW sedge{sedge_bs->p, sedge_bs->len};
    std::vector<double> sl((size_t)nlev * L, 0.0), sr((size_t)nlev * L, 0.0);
    RowW sedgel{sl.data(), L}, sedger{sr.data(), L};

    for (int n = 0; n < nlev; ++n) {
        const int nr = rg->nr[n];
        const double dr = rg->dr[n];
        const double dtdr = dt / dr;
        for (int i = 0; i < rg->nchunks[n]; ++i) {
            const int lo = rg->r_start[n][i];
            const int hi = rg->r_end[n][i];
            if (p.ppm_type == 0) {
                for (int r = lo; r <= hi; ++r) {
                    double slope = 0.0;
                    if (p.slope_order == 0) {
                        slope = 0.0;
                    } else if (p.slope_order == 2) {
                        if (r == 0) {
                            slope = s(n, r + 1) - s(n, r);
                        } else if (r == nr - 1) {
                            slope = s(n, r) - s(n, r - 1);
                        } else {
                            double del = 0.5 * (s(n, r + 1) - s(n, r - 1));
                            double dpls = 2.0 * (s(n, r + 1) - s(n, r));
                            double dmin = 2.0 * (s(n, r) - s(n, r - 1));
                            double slim = mn(ab(dpls), ab(dmin));
                            slim = dpls * dmin > 0.0 ? slim : 0.0;
                            double sflag = cs(1.0, del);
                            slope = sflag * mn(slim, ab(del));
                        }
                    } else if (p.slope_order == 4) {
                        double dxm = fromm_planar(s, n, r - 1, nr);
                        double dxp = fromm_planar(s, n, r + 1, nr);
                        if (r == 0) {
                            slope = s(n, r + 1) - s(n, r);
                        } else if (r == nr - 1) {
                            slope = s(n, r) - s(n, r - 1);
                        } else {
                            double cen = 0.5 * (s(n, r + 1) - s(n, r - 1));
                            double dpls = 2.0 * (s(n, r + 1) - s(n, r));
                            double dmin = 2.0 * (s(n, r) - s(n, r - 1));
                            double lim = mn(ab(dmin), ab(dpls));
                            lim = dpls * dmin > 0.0 ? lim : 0.0;
                            double flag = cs(1.0, cen);
                            double ds = 4.0 / 3.0 * cen - (dxp + dxm) / 6.0;
                            slope = flag * mn(ab(ds), lim);
                        }
                    }
                    double u = 0.5 * (w0(n, r) + w0(n, r + 1));
                    double ubardth = dth * u / dr;
                    sedgel(n, r + 1) = s(n, r) + (0.5 - ubardth) * slope + dth * force(n, r);
                    sedger(n, r) = s(n, r) - (0.5 + ubardth) * slope + dth * force(n, r);
                }
                continue;
            }
            std::vector<double> sedget(rg->nr_fine + 1, 0.0);
            if (p.ppm_type == 2) {
                const double C = 1.25;
                for (int r = lo; r <= hi + 1; ++r) {
                    double dsvl = 0.0;
                    if (r - 1 == 0) {
                        dsvl = s(n, r) - s(n, r - 1);
                    } else if (r - 1 == nr - 1) {
                        dsvl = s(n, r - 1) - s(n, r - 2);
                    } else if (r - 1 > 0 && r - 1 < nr - 1) {
                        dsvl = 0.5 * (s(n, r) - s(n, r - 2));
                    }
                    double dsvr = 0.0;
                    if (r == 0) {
                        dsvr = s(n, r + 1) - s(n, r);
                    } else if (r == nr - 1) {
                        dsvr = s(n, r) - s(n, r - 1);
                    } else if (r > 0 && r < nr - 1) {
                        dsvr = 0.5 * (s(n, r + 1) - s(n, r - 1));
                    }
                    if (r == 0) {
                        sedget[r] = s(n, r) - 0.5 * dsvr;
                    } else if (r == nr) {
                        sedget[r] = s(n, r - 1) + 0.5 * dsvr;
                    } else if (r > 0 && r < nr) {
                        sedget[r] = 0.5 * (s(n, r) + s(n, r - 1)) - (dsvr - dsvl) / 6.0;
                        if (r >= 2 && r <= nr - 2) {
                            if ((sedget[r] - s(n, r - 1)) * (s(n, r) - sedget[r]) < 0.0) {
                                double D2 = 3.0 * (s(n, r - 1) - 2.0 * sedget[r] + s(n, r));
                                double D2L = s(n, r - 2) - 2.0 * s(n, r - 1) + s(n, r);
                                double D2R = s(n, r - 1) - 2.0 * s(n, r) + s(n, r + 1);
                                double sgn = cs(1.0, D2);
                                double D2LIM =
                                    sgn * mx(mn(C * sgn * D2L, mn(C * sgn * D2R, sgn * D2)), 0.0);
                                sedget[r] = 0.5 * (s(n, r - 1) + s(n, r)) - D2LIM / 6.0;
                            }
                        }
                    }
                }
            }
            for (int r = lo; r <= hi; ++r) {
                double sm = 0.0, sp = 0.0;
                if (p.ppm_type == 1) {
                    double dsvlm = vl_planar(s, n, r - 1, nr);
                    double dsvl = vl_planar(s, n, r, nr);
                    if (r == 0) {
                        sm = s(n, r) - 0.5 * dsvl;
                    } else if (r == nr) {
                        sm = s(n, r - 1) + 0.5 * dsvl;
                    } else {
                        sm = 0.5 * (s(n, r) + s(n, r - 1)) - (dsvl - dsvlm) / 6.0;
                        sm = mx(sm, mn(s(n, r), s(n, r - 1)));
                        sm = mn(sm, mx(s(n, r), s(n, r - 1)));
                    }
                    dsvlm = dsvl;
                    const int rp = r + 1;
                    dsvl = vl_planar(s, n, rp, nr);
                    if (rp == 0) {
                        sp = s(n, rp) - 0.5 * dsvl;
                    } else if (rp == nr) {
                        sp = s(n, rp - 1) + 0.5 * dsvl;
                    } else {
                        sp = 0.5 * (s(n, rp) + s(n, rp - 1)) - (dsvl - dsvlm) / 6.0;
                        sp = mx(sp, mn(s(n, rp), s(n, rp - 1)));
                        sp = mn(sp, mx(s(n, rp), s(n, rp - 1)));
                    }
                    cw84(s(n, r), sm, sp);
                } else {
                    if (r >= 2 && r <= nr - 3) {
                        cs08(s(n, r - 2), s(n, r - 1), s(n, r), s(n, r + 1), s(n, r + 2),
                             sedget[r - 1], sedget[r], sedget[r + 1], sedget[r + 2], sm, sp);
                    } else {
                        sp = sedget[r + 1];
                        sm = sedget[r];
                    }
                }
                double sigmap = ab(w0(n, r + 1)) * dtdr;
                double sigmam = ab(w0(n, r)) * dtdr;
                double s6 = 6.0 * s(n, r) - 3.0 * (sm + sp);
                double Ip = 0.0, Im = 0.0;
                if (w0(n, r + 1) > rel_eps) {
                    if (p.ppm_type == 1)
                        Ip = sp - (sigmap / 2.0) * (sp - sm - (1.0 - 2.0 / 3.0 * sigmap) * s6);
                    else
                        Ip = sp - 0.5 * sigmap * (sp - sm - (1.0 - 2.0 / 3.0 * sigmap) * s6);
                } else {
                    Ip = s(n, r);
                }
                if (w0(n, r) < -rel_eps) {
                    if (p.ppm_type == 1)
                        Im = sm + (sigmam / 2.0) * (sp - sm + (1.0 - 2.0 / 3.0 * sigmam) * s6);
                    else
                        Im = sm + 0.5 * sigmam * (sp - sm + (1.0 - 2.0 / 3.0 * sigmam) * s6);
                } else {
                    Im = s(n, r);
                }
                sedgel(n, r + 1) = Ip + dth * force(n, r);
                sedger(n, r) = Im + dth * force(n, r);
            }
        }
    }

    // coarse-fine synchronisation (:927-950)
    for (int n = 0; n < nlev; ++n) {
        for (int i = 0; i < rg->nchunks[n]; ++i) {
            const int lo = rg->r_start[n][i];
            const int hi = rg->r_end[n][i];
            if (n < nlev - 1) {
                sedger(n, rg->r_start[n + 1][i] / 2) = sedger(n + 1, rg->r_start[n + 1][i]);
                sedgel(n, (rg->r_end[n + 1][i] + 1) / 2) = sedgel(n + 1, rg->r_end[n + 1][i] + 1);
            }
            if (n > 0) {
                sedgel(n, lo) = sedgel(n - 1, lo / 2);
                sedger(n, hi + 1) = sedger(n - 1, (hi + 1) / 2);
            }
        }
    }

    // Riemann problem (:952-979)
    for (int n = 0; n < nlev; ++n) {
        const int nr = rg->nr[n];
        for (int i = 0; i < rg->nchunks[n]; ++i) {
            const int lo = rg->r_start[n][i];
            const int hi = rg->r_end[n][i];
            for (int r = lo; r <= hi + 1; ++r) {
                if (r == 0) {
                    sedge(n, r) = sedger(n, r);
                } else if (r == nr) {
                    sedge(n, r) = sedgel(n, r);
                } else {
                    sedge(n, r) = w0(n, r) > 0.0 ? sedgel(n, r) : sedger(n, r);
                    sedge(n, r) =
                        ab(w0(n, r)) < rel_eps ? 0.5 * (sedger(n, r) + sedgel(n, r)) : sedge(n, r);
                }
            }
        }
    }
}

}  // namespace

// Maestro::MakeEdgeState1d (MaestroMakeEdgeState.cpp:5-15)
int MakeEdgeState1d(const Ctx* c, const mx_radial_geom* rg, const mx_basestate* s,
                    mx_basestate* sedge, const mx_basestate* force, const mx_basestate* w0) {
    if (c->p.spherical)
        edge1d_sphr(c, rg, s, sedge, force, w0);
    else
        edge1d_planar(c, rg, s, sedge, force, w0);
    return MX_OK;
}

}  // namespace mxo
