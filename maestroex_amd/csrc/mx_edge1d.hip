// mx_edge1d.hip -- Maestro::MakeEdgeState1d on the GPU.
//
// The base state is one-dimensional (nr_fine <= a few thousand cells), so the
// whole function runs in one workgroup: each phase of the reference
// (MaestroMakeEdgeState.cpp) is a loop over radial cells spread across the
// block's threads, and phases are separated by barriers.  Temporaries
// (sedgel, sedger, the type-2 face values) live in the context workspace and
// are cleared by the kernel, matching the zero-filled BaseState temporaries
// of the reference (BaseState.H:308-316).
#include "mx_internal.H"

using namespace mxd;
using namespace mxh;

namespace {

constexpr int NT1 = 512;

struct E1Args {
    const double* s;
    const double* f;
    const double* w0;
    double* sedge;
    long long slen, flen, wlen, elen;
    double* sl;   // sedgel, nlev x (nr_fine + 1)
    double* sr;   // sedger
    double* sgt;  // type-2 face values, nr_fine + 1
    int ppm_type, slope_order, spherical;
    double dt, rel_eps;
};

// van Leer slope of the spherical array with three reflected (centre) and
// extrapolated (top) ghost cells (MaestroMakeEdgeState.cpp:44-57)
__device__ __forceinline__ double sg_at(const E1Args& A, int nr, int k) {
    const int ng = 3;
    if (k < ng) return A.s[ng - 1 - k];
    if (k >= ng + nr) return A.s[nr - 1];
    return A.s[k - ng];
}

__device__ __forceinline__ void cw84(double s0, double& sm, double& sp) {
    if ((sp - s0) * (s0 - sm) <= 0.0) {
        sp = s0;
        sm = s0;
    } else if (dabs(sp - s0) >= 2.0 * dabs(sm - s0)) {
        sp = 3.0 * s0 - 2.0 * sm;
    } else if (dabs(sm - s0) >= 2.0 * dabs(sp - s0)) {
        sm = 3.0 * s0 - 2.0 * sp;
    }
}

// limited van Leer slope of planar level data at rr, one-sided at the ends
__device__ __forceinline__ double vl_planar(const double* s, int rr, int nr) {
    double d = 0.0;
    if (rr == 0) {
        d = s[rr + 1] - s[rr];
    } else if (rr == nr - 1) {
        d = s[rr] - s[rr - 1];
    } else if (rr > 0 && rr < nr - 1) {
        d = vl_slope(s[rr - 1], s[rr], s[rr + 1]);
    }
    return d;
}

// limited Fromm slope of planar level data at rr (slope_order 4 neighbours)
__device__ __forceinline__ double fromm_planar(const double* s, int rr, int nr) {
    double d = 0.0;
    if (rr == 0) {
        d = s[rr + 1] - s[rr];
    } else if (rr == nr - 1) {
        d = s[rr] - s[rr - 1];
    } else if (rr > 0 && rr < nr - 1) {
        double cen = 0.5 * (s[rr + 1] - s[rr - 1]);
        double dpls = 2.0 * (s[rr + 1] - s[rr]);
        double dmn = 2.0 * (s[rr] - s[rr - 1]);
        double lim = dmin(dabs(dmn), dabs(dpls));
        lim = dpls * dmn > 0.0 ? lim : 0.0;
        double flag = dsign1(cen);
        d = flag * dmin(lim, dabs(cen));
    }
    return d;
}

// spherical edge states of cell r (:59-330)
__device__ void sphr_cell(const E1Args& A, int nr, double dr0, int r) {
    const double dth = 0.5 * A.dt;
    const double dtdr = A.dt / dr0;
    const int ng = 3;
    const double s0 = A.s[r];
    const double fr = A.f[r];
    const double wlo = A.w0[r], whi = A.w0[r + 1];
    if (A.ppm_type == 0) {
        double slope = 0.0;
        if (A.slope_order == 2) {
            const int g = r + ng;
            const double a = sg_at(A, nr, g - 1), b = sg_at(A, nr, g), c = sg_at(A, nr, g + 1);
            double del = 0.5 * (c - a);
            double dpls = 2.0 * (c - b);
            double dmn = 2.0 * (b - a);
            double slim = dmin(dabs(dpls), dabs(dmn));
            slim = dpls * dmn > 0.0 ? slim : 0.0;
            double sflag = dsign1(del);
            slope = sflag * dmin(slim, dabs(del));
        } else if (A.slope_order == 4) {
            double cen[3], lim[3], flag[3], fromm[3];
            for (int i = 0; i < 3; ++i) {
                const int g = r + ng - 1 + i;
                const double a = sg_at(A, nr, g - 1), b = sg_at(A, nr, g),
                             c = sg_at(A, nr, g + 1);
                cen[i] = 0.5 * (c - a);
                double dpls = 2.0 * (c - b);
                double dmn = 2.0 * (b - a);
                lim[i] = dmin(dabs(dmn), dabs(dpls));
                lim[i] = dpls * dmn > 0.0 ? lim[i] : 0.0;
                flag[i] = dsign1(cen[i]);
                fromm[i] = flag[i] * dmin(lim[i], dabs(cen[i]));
            }
            double ds = 4.0 / 3.0 * cen[1] - (fromm[2] + fromm[0]) / 6.0;
            slope = flag[1] * dmin(dabs(ds), lim[1]);
        }
        double u = 0.5 * (wlo + whi);
        double ubardth = dth * u / dr0;
        A.sl[r + 1] = s0 + (0.5 - ubardth) * slope + dth * fr;
        A.sr[r] = s0 - (0.5 + ubardth) * slope + dth * fr;
        return;
    }
    double v[9];
    for (int o = -4; o <= 4; ++o) v[4 + o] = sg_at(A, nr, r + ng + o);
    double sm, sp;
    if (A.ppm_type == 1) {
        const double dl = vl_slope(v[2], v[3], v[4]);
        const double dc = vl_slope(v[3], v[4], v[5]);
        const double dr = vl_slope(v[4], v[5], v[6]);
        sm = 0.5 * (v[4] + v[3]) - (dc - dl) / 6.0;
        sm = dmax(sm, dmin(v[4], v[3]));
        sm = dmin(sm, dmax(v[4], v[3]));
        sp = 0.5 * (v[5] + v[4]) - (dr - dc) / 6.0;
        sp = dmax(sp, dmin(v[5], v[4]));
        sp = dmin(sp, dmax(v[5], v[4]));
        cw84(s0, sm, sp);
    } else {
        const double f0 = cs_face(v[1], v[2], v[3], v[4]);
        const double f1 = cs_face(v[2], v[3], v[4], v[5]);
        const double f2 = cs_face(v[3], v[4], v[5], v[6]);
        const double f3 = cs_face(v[4], v[5], v[6], v[7]);
        cs_cell(v, f0, f1, f2, f3, 1.e-10, sm, sp);
    }
    double sigmap = dabs(whi) * dtdr;
    double sigmam = dabs(wlo) * dtdr;
    double s6 = 6.0 * s0 - 3.0 * (sm + sp);
    double Ip = whi > A.rel_eps ? sp - 0.5 * sigmap * (sp - sm - (1.0 - 2.0 / 3.0 * sigmap) * s6)
                                : s0;
    double Im = wlo < -A.rel_eps ? sm + 0.5 * sigmam * (sp - sm + (1.0 - 2.0 / 3.0 * sigmam) * s6)
                                 : s0;
    A.sl[r + 1] = Ip + dth * fr;
    A.sr[r] = Im + dth * fr;
}

// planar type-2 face value r of level n (:722-780)
__device__ double planar_face2(const double* s, int r, int nr) {
    const double C = 1.25;
    double dsvl = 0.0;
    if (r - 1 == 0) {
        dsvl = s[r] - s[r - 1];
    } else if (r - 1 == nr - 1) {
        dsvl = s[r - 1] - s[r - 2];
    } else if (r - 1 > 0 && r - 1 < nr - 1) {
        dsvl = 0.5 * (s[r] - s[r - 2]);
    }
    double dsvr = 0.0;
    if (r == 0) {
        dsvr = s[r + 1] - s[r];
    } else if (r == nr - 1) {
        dsvr = s[r] - s[r - 1];
    } else if (r > 0 && r < nr - 1) {
        dsvr = 0.5 * (s[r + 1] - s[r - 1]);
    }
    double e = 0.0;
    if (r == 0) {
        e = s[r] - 0.5 * dsvr;
    } else if (r == nr) {
        e = s[r - 1] + 0.5 * dsvr;
    } else if (r > 0 && r < nr) {
        e = 0.5 * (s[r] + s[r - 1]) - (dsvr - dsvl) / 6.0;
        if (r >= 2 && r <= nr - 2) {
            if ((e - s[r - 1]) * (s[r] - e) < 0.0) {
                double D2 = 3.0 * (s[r - 1] - 2.0 * e + s[r]);
                double D2L = s[r - 2] - 2.0 * s[r - 1] + s[r];
                double D2R = s[r - 1] - 2.0 * s[r] + s[r + 1];
                double sgn = dsign1(D2);
                double D2LIM = sgn * dmax(dmin(C * sgn * D2L, dmin(C * sgn * D2R, sgn * D2)), 0.0);
                e = 0.5 * (s[r - 1] + s[r]) - D2LIM / 6.0;
            }
        }
    }
    return e;
}

// planar edge states of cell r of level n (:439-919)
__device__ void planar_cell(const E1Args& A, const double* s, const double* f, const double* w0,
                            double* sl, double* sr, int nr, double dr, int r) {
    const double dth = 0.5 * A.dt;
    const double dtdr = A.dt / dr;
    if (A.ppm_type == 0) {
        double slope = 0.0;
        if (A.slope_order == 2) {
            if (r == 0) {
                slope = s[r + 1] - s[r];
            } else if (r == nr - 1) {
                slope = s[r] - s[r - 1];
            } else {
                double del = 0.5 * (s[r + 1] - s[r - 1]);
                double dpls = 2.0 * (s[r + 1] - s[r]);
                double dmn = 2.0 * (s[r] - s[r - 1]);
                double slim = dmin(dabs(dpls), dabs(dmn));
                slim = dpls * dmn > 0.0 ? slim : 0.0;
                double sflag = dsign1(del);
                slope = sflag * dmin(slim, dabs(del));
            }
        } else if (A.slope_order == 4) {
            double dxm = fromm_planar(s, r - 1, nr);
            double dxp = fromm_planar(s, r + 1, nr);
            if (r == 0) {
                slope = s[r + 1] - s[r];
            } else if (r == nr - 1) {
                slope = s[r] - s[r - 1];
            } else {
                double cen = 0.5 * (s[r + 1] - s[r - 1]);
                double dpls = 2.0 * (s[r + 1] - s[r]);
                double dmn = 2.0 * (s[r] - s[r - 1]);
                double lim = dmin(dabs(dmn), dabs(dpls));
                lim = dpls * dmn > 0.0 ? lim : 0.0;
                double flag = dsign1(cen);
                double ds = 4.0 / 3.0 * cen - (dxp + dxm) / 6.0;
                slope = flag * dmin(dabs(ds), lim);
            }
        }
        double u = 0.5 * (w0[r] + w0[r + 1]);
        double ubardth = dth * u / dr;
        sl[r + 1] = s[r] + (0.5 - ubardth) * slope + dth * f[r];
        sr[r] = s[r] - (0.5 + ubardth) * slope + dth * f[r];
        return;
    }
    double sm = 0.0, sp = 0.0;
    if (A.ppm_type == 1) {
        double dsvlm = vl_planar(s, r - 1, nr);
        double dsvl = vl_planar(s, r, nr);
        if (r == 0) {
            sm = s[r] - 0.5 * dsvl;
        } else {
            sm = 0.5 * (s[r] + s[r - 1]) - (dsvl - dsvlm) / 6.0;
            sm = dmax(sm, dmin(s[r], s[r - 1]));
            sm = dmin(sm, dmax(s[r], s[r - 1]));
        }
        dsvlm = dsvl;
        const int rp = r + 1;
        dsvl = vl_planar(s, rp, nr);
        if (rp == nr) {
            sp = s[rp - 1] + 0.5 * dsvl;
        } else {
            sp = 0.5 * (s[rp] + s[rp - 1]) - (dsvl - dsvlm) / 6.0;
            sp = dmax(sp, dmin(s[rp], s[rp - 1]));
            sp = dmin(sp, dmax(s[rp], s[rp - 1]));
        }
        cw84(s[r], sm, sp);
    } else {
        const double* g = A.sgt;
        if (r >= 2 && r <= nr - 3) {
            double v[9];
            for (int o = -2; o <= 2; ++o) v[4 + o] = s[r + o];
            v[0] = v[1] = v[7] = v[8] = 0.0;
            cs_cell(v, g[r - 1], g[r], g[r + 1], g[r + 2], 1.e-10, sm, sp);
        } else {
            sp = g[r + 1];
            sm = g[r];
        }
    }
    double sigmap = dabs(w0[r + 1]) * dtdr;
    double sigmam = dabs(w0[r]) * dtdr;
    double s6 = 6.0 * s[r] - 3.0 * (sm + sp);
    double Ip, Im;
    if (w0[r + 1] > A.rel_eps) {
        if (A.ppm_type == 1)
            Ip = sp - (sigmap / 2.0) * (sp - sm - (1.0 - 2.0 / 3.0 * sigmap) * s6);
        else
            Ip = sp - 0.5 * sigmap * (sp - sm - (1.0 - 2.0 / 3.0 * sigmap) * s6);
    } else {
        Ip = s[r];
    }
    if (w0[r] < -A.rel_eps) {
        if (A.ppm_type == 1)
            Im = sm + (sigmam / 2.0) * (sp - sm + (1.0 - 2.0 / 3.0 * sigmam) * s6);
        else
            Im = sm + 0.5 * sigmam * (sp - sm + (1.0 - 2.0 / 3.0 * sigmam) * s6);
    } else {
        Im = s[r];
    }
    sl[r + 1] = Ip + dth * f[r];
    sr[r] = Im + dth * f[r];
}

__global__ void __launch_bounds__(NT1) k_edge_state_1d(mx_radial_geom rg, E1Args A) {
    const int t = threadIdx.x;
    const int L = rg.nr_fine + 1;
    const int nlev = A.spherical ? 1 : rg.nlev;
    for (int i = t; i < nlev * L; i += NT1) {
        A.sl[i] = 0.0;
        A.sr[i] = 0.0;
    }
    __syncthreads();

    if (A.spherical) {
        const int nr = rg.nr_fine;
        for (int r = t; r < nr; r += NT1) sphr_cell(A, nr, rg.dr[0], r);
        __syncthreads();
        for (int r = t; r <= nr; r += NT1) {
            double l = A.sl[r], rr = A.sr[r];
            if (r == 0) {
                l = rr;
            } else if (r == nr) {
                rr = l;
            }
            const double w = A.w0[r];
            double e = w > 0.0 ? l : rr;
            e = dabs(w) < A.rel_eps ? 0.5 * (rr + l) : e;
            A.sedge[r] = e;
        }
        return;
    }

    for (int n = 0; n < nlev; ++n) {
        const int nr = rg.nr[n];
        const double* s = A.s + (long long)n * A.slen;
        const double* f = A.f + (long long)n * A.flen;
        const double* w0 = A.w0 + (long long)n * A.wlen;
        double* sl = A.sl + (long long)n * L;
        double* sr = A.sr + (long long)n * L;
        for (int i = 0; i < rg.nchunks[n]; ++i) {
            const int lo = rg.r_start[n][i];
            const int hi = rg.r_end[n][i];
            if (A.ppm_type == 2) {
                for (int r = t; r < L; r += NT1) A.sgt[r] = 0.0;
                __syncthreads();
                for (int r = lo + t; r <= hi + 1; r += NT1) A.sgt[r] = planar_face2(s, r, nr);
                __syncthreads();
            }
            for (int r = lo + t; r <= hi; r += NT1) planar_cell(A, s, f, w0, sl, sr, nr, rg.dr[n], r);
            __syncthreads();
        }
    }
    // coarse-fine synchronisation, in the reference's order (:927-950)
    if (t == 0) {
        for (int n = 0; n < nlev; ++n) {
            for (int i = 0; i < rg.nchunks[n]; ++i) {
                const int lo = rg.r_start[n][i];
                const int hi = rg.r_end[n][i];
                if (n < nlev - 1) {
                    A.sr[n * L + rg.r_start[n + 1][i] / 2] = A.sr[(n + 1) * L + rg.r_start[n + 1][i]];
                    A.sl[n * L + (rg.r_end[n + 1][i] + 1) / 2] =
                        A.sl[(n + 1) * L + rg.r_end[n + 1][i] + 1];
                }
                if (n > 0) {
                    A.sl[n * L + lo] = A.sl[(n - 1) * L + lo / 2];
                    A.sr[n * L + hi + 1] = A.sr[(n - 1) * L + (hi + 1) / 2];
                }
            }
        }
    }
    __syncthreads();
    // Riemann problem (:952-979)
    for (int n = 0; n < nlev; ++n) {
        const int nr = rg.nr[n];
        const double* w0 = A.w0 + (long long)n * A.wlen;
        double* e = A.sedge + (long long)n * A.elen;
        const double* sl = A.sl + (long long)n * L;
        const double* sr = A.sr + (long long)n * L;
        for (int i = 0; i < rg.nchunks[n]; ++i) {
            const int lo = rg.r_start[n][i];
            const int hi = rg.r_end[n][i];
            for (int r = lo + t; r <= hi + 1; r += NT1) {
                double v;
                if (r == 0) {
                    v = sr[r];
                } else if (r == nr) {
                    v = sl[r];
                } else {
                    v = w0[r] > 0.0 ? sl[r] : sr[r];
                    v = dabs(w0[r]) < A.rel_eps ? 0.5 * (sr[r] + sl[r]) : v;
                }
                e[r] = v;
            }
        }
    }
}

}  // namespace

extern "C" int mx_adv_make_edge_state_1d(mx_adv_ctx* c, const mx_radial_geom* rg,
                                         const mx_basestate* s, mx_basestate* sedge,
                                         const mx_basestate* force, const mx_basestate* w0) {
    if (!c || !rg || !s || !sedge || !force || !w0) return MX_ERR_ARG;
    if (!s->p || !sedge->p || !force->p || !w0->p) return MX_ERR_ARG;
    const bool sph = c->p.spherical != 0;
    const int nlev = sph ? 1 : rg->nlev;
    if (nlev < 1 || nlev > MX_MAX_LEV || rg->nr_fine < 4) return MX_ERR_ARG;
    if (s->len < rg->nr_fine || force->len < rg->nr_fine || sedge->len < rg->nr_fine + 1 ||
        w0->len < rg->nr_fine + 1)
        return MX_ERR_ARG;
    if (s->nlev < nlev || force->nlev < nlev || sedge->nlev < nlev || w0->nlev < nlev)
        return MX_ERR_ARG;
    if (sph) {
        if (!(rg->dr[0] > 0.0)) return MX_ERR_ARG;
    } else {
        for (int n = 0; n < nlev; ++n) {
            const int nr = rg->nr[n];
            if (nr < 4 || nr > rg->nr_fine || !(rg->dr[n] > 0.0)) return MX_ERR_ARG;
            if (rg->nchunks[n] < 0 || rg->nchunks[n] > MX_MAX_CHUNK) return MX_ERR_ARG;
            if (n + 1 < nlev && rg->nchunks[n + 1] < rg->nchunks[n]) return MX_ERR_UNSUPPORTED;
            for (int i = 0; i < rg->nchunks[n]; ++i) {
                const int lo = rg->r_start[n][i], hi = rg->r_end[n][i];
                if (lo < 0 || hi < lo || hi > nr - 1) return MX_ERR_ARG;
                // "make_edge_state assumes blocking_factor > 1" (MaestroMakeEdgeState.cpp:427-436)
                if ((c->p.ppm_type == 1 || c->p.ppm_type == 2) && (lo == 2 || hi == nr - 3))
                    return MX_ERR_ARG;
                if (n > 0 && (lo / 2 > rg->nr[n - 1] || (hi + 1) / 2 > rg->nr[n - 1]))
                    return MX_ERR_ARG;
            }
        }
    }
    const size_t L = (size_t)rg->nr_fine + 1;
    int rc = ws_reserve(c, (2 * (size_t)nlev * L + L) * sizeof(double) + 1024);
    if (rc) return rc;
    E1Args A;
    A.s = s->p;
    A.f = force->p;
    A.w0 = w0->p;
    A.sedge = sedge->p;
    A.slen = s->len;
    A.flen = force->len;
    A.wlen = w0->len;
    A.elen = sedge->len;
    A.sl = ws_take(c, (size_t)nlev * L);
    A.sr = ws_take(c, (size_t)nlev * L);
    A.sgt = ws_take(c, L);
    A.ppm_type = c->p.ppm_type;
    A.slope_order = c->p.slope_order;
    A.spherical = sph ? 1 : 0;
    A.dt = c->p.dt;
    A.rel_eps = c->p.rel_eps;
    MX_LAUNCH(c, "k_edge_state_1d", k_edge_state_1d, dim3(1), dim3(NT1), *rg, A);
    return MX_OK;
}
