// mx_device.H -- device-side building blocks of the advection kernels.
//
// Array4-like views, the min/max/abs/copysign semantics of the reference
// (amrex::min/max are std::min/std::max: ties return the first argument),
// and per-point reconstruction functions: the PPM edge values and
// characteristic trace (Source/MaestroPPM.cpp) and the limited slopes
// (Source/MaestroSlopes.cpp), written once for a generic direction with the
// reference's per-direction differences selected at compile time.
//
// All arithmetic is FP64 in the operand order of the reference expressions;
// the library is compiled without FMA contraction so every result is the
// correctly rounded IEEE value the reference's CPU build produces.
#ifndef MX_DEVICE_H
#define MX_DEVICE_H

#include <hip/hip_runtime.h>

#include "mx_adv.h"

namespace mxd {

struct DA {
    double* p;
    long long js, ks, ns;
    int b0, b1, b2;
    int pad_ = 0;  // explicit, so that copies of a view compare equal byte for byte
    __device__ __forceinline__ double& operator()(int i, int j, int k, int n = 0) const {
        return p[(long long)(i - b0) + (long long)(j - b1) * js + (long long)(k - b2) * ks +
                 (long long)n * ns];
    }
    __device__ __forceinline__ long long off(int i, int j, int k, int n = 0) const {
        return (long long)(i - b0) + (long long)(j - b1) * js + (long long)(k - b2) * ks +
               (long long)n * ns;
    }
    // stride of one step in direction d
    __device__ __forceinline__ long long stride(int d) const {
        return d == 0 ? 1 : (d == 1 ? js : ks);
    }
};

__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double dabs(double a) { return fabs(a); }
// min / max of two values that are both >= +0 and not NaN: every such pair
// orders totally and equal values have equal bits, so the hardware min / max
// (one instruction) returns what dmin / dmax return
__device__ __forceinline__ double nn_min(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ double nn_max(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ double dsign1(double b) { return copysign(1.0, b); }

// ---------------------------------------------------------------------------
// PPM edge values of one cell along direction D (MaestroPPM.cpp).
// `q` points at s(cell), `st` is the stride of one cell along D.

__device__ __forceinline__ double vl_slope(double a, double b, double c) {
    double dsc = 0.5 * (c - a);
    double dsl = 2.0 * (b - a);
    double dsr = 2.0 * (c - b);
    double r = 0.0;
    // sign(dsc) * m with m >= +0 is m with the sign bit of dsc
    if (dsl * dsr > 0.0) r = copysign(nn_min(dabs(dsc), nn_min(dabs(dsl), dabs(dsr))), dsc);
    return r;
}

__device__ __forceinline__ void cw_limit(double s0, double& sm, double& sp) {
    if ((sp - s0) * (s0 - sm) <= 0.0) {
        sp = s0;
        sm = s0;
    } else if (dabs(sp - s0) >= 2.0 * dabs(sm - s0)) {
        sp = 3.0 * s0 - 2.0 * sm;
    } else if (dabs(sm - s0) >= 2.0 * dabs(sp - s0)) {
        sm = 3.0 * s0 - 2.0 * sp;
    }
}

__device__ __forceinline__ double bnd4(double a, double b, double c, double d) {
    return -0.2 * a + 0.75 * b + 0.5 * c - 0.05 * d;
}

__device__ __forceinline__ double cs_face(double sa, double sb, double sc, double sd) {
    const double C = 1.25;
    double e = (7.0 / 12.0) * (sb + sc) - (1.0 / 12.0) * (sa + sd);
    if ((e - sb) * (sc - e) < 0.0) {
        double D2 = 3.0 * (sb - 2.0 * e + sc);
        double D2L = sa - 2.0 * sb + sc;
        double D2R = sb - 2.0 * sc + sd;
        double sgn = dsign1(D2);
        double D2LIM = sgn * dmax(dmin(C * sgn * D2L, dmin(C * sgn * D2R, sgn * D2)), 0.0);
        e = 0.5 * (sb + sc) - D2LIM / 6.0;
    }
    return e;
}

// cell limiter of the Colella-Sekora scheme; v[o+4] = s(cell + o)
__device__ __forceinline__ void cs_cell(const double* v, double sedgel, double sedge,
                                        double sedger, double sedgerr, double bigm_thresh,
                                        double& sm, double& sp) {
    const double C = 1.25;
    const double s0 = v[4];
    double alphap = sedger - s0;
    double alpham = sedge - s0;
    bool bigp = dabs(alphap) > 2.0 * dabs(alpham);
    bool bigm = dabs(alpham) > 2.0 * dabs(alphap);
    bool extremum = false;
    if (alpham * alphap >= 0.0) {
        extremum = true;
    } else if (bigp || bigm) {
        double dafacem = sedge - sedgel;
        double dafacep = sedgerr - sedger;
        double dabarm = s0 - v[3];
        double dabarp = v[5] - s0;
        double dafacemin = nn_min(dabs(dafacem), dabs(dafacep));
        double dabarmin = nn_min(dabs(dabarm), dabs(dabarp));
        double dachkm, dachkp;
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
        double D2L = v[2] - 2.0 * v[3] + s0;
        double D2R = s0 - 2.0 * v[5] + v[6];
        double D2C = v[3] - 2.0 * s0 + v[5];
        double sgn = dsign1(D2);
        double D2LIM =
            dmax(dmin(sgn * D2, dmin(C * sgn * D2L, dmin(C * sgn * D2R, C * sgn * D2C))), 0.0);
        double D2ABS = nn_max(dabs(D2), 1.e-10);
        alpham = alpham * D2LIM / D2ABS;
        alphap = alphap * D2LIM / D2ABS;
    } else {
        if (bigp) {
            double sgn = dsign1(alpham);
            double amax = -alphap * alphap / (4.0 * (alpham + alphap));
            double delam = v[3] - s0;
            if (sgn * amax >= sgn * delam) {
                if (sgn * (delam - alpham) >= 1.e-10) {
                    alphap = -2.0 * delam - 2.0 * sgn * sqrt(delam * delam - delam * alpham);
                } else {
                    alphap = -2.0 * alpham;
                }
            }
        }
        if (bigm) {
            double sgn = dsign1(alphap);
            double amax = -alpham * alpham / (4.0 * (alpham + alphap));
            double delap = v[5] - s0;
            if (sgn * amax >= sgn * delap) {
                if (sgn * (delap - alphap) >= bigm_thresh) {
                    alpham = -2.0 * delap - 2.0 * sgn * sqrt(delap * delap - delap * alphap);
                } else {
                    alpham = -2.0 * alphap;
                }
            }
        }
    }
    sm = s0 + alpham;
    sp = s0 + alphap;
}

__device__ __forceinline__ bool is_ho(int bc) { return bc == MX_BC_EXT_DIR || bc == MX_BC_HOEXTRAP; }

// Edge values sm, sp of the cell whose stencil is v[0..8] = s(cell-4 .. cell+4)
// (ppm_type 1 reads v[2..6], ppm_type 2 v[1..7]).
template <int PPM, int D>
__device__ __forceinline__ void ppm_edges(const double* v, int pos, int dlo, int dhi, int bclo,
                                          int bchi, double& sm_out, double& sp_out) {
    double sm, sp;
    if (PPM == 1) {
        double dl = vl_slope(v[2], v[3], v[4]);
        double dc = vl_slope(v[3], v[4], v[5]);
        double dr = vl_slope(v[4], v[5], v[6]);
        sm = 0.5 * (v[4] + v[3]) - (dc - dl) / 6.0;
        sm = dmax(sm, dmin(v[4], v[3]));
        sm = dmin(sm, dmax(v[4], v[3]));
        sp = 0.5 * (v[5] + v[4]) - (dr - dc) / 6.0;
        sp = dmax(sp, dmin(v[5], v[4]));
        sp = dmin(sp, dmax(v[5], v[4]));
        const double sedgel = sp;
        const double sedger = sm;
        cw_limit(v[4], sm, sp);
        if (pos == dlo) {
            if (is_ho(bclo)) {
                sm = v[3];
                sp = bnd4(v[3], v[4], v[5], v[6]);
                sp = dmax(sp, dmin(v[5], v[4]));
                sp = dmin(sp, dmax(v[5], v[4]));
            }
        } else if (pos == dlo + 1) {
            if (is_ho(bclo)) {
                sm = bnd4(v[2], v[3], v[4], v[5]);
                sm = dmax(sm, dmin(v[4], v[3]));
                sm = dmin(sm, dmax(v[4], v[3]));
                sp = sedgel;
                cw_limit(v[4], sm, sp);
            }
        } else if (pos == dhi) {
            if (is_ho(bchi)) {
                sp = v[5];
                sm = bnd4(v[5], v[4], v[3], v[2]);
                sm = dmax(sm, dmin(v[3], v[4]));
                sm = dmin(sm, dmax(v[3], v[4]));
            }
        } else if (pos == dhi - 1) {
            if (is_ho(bchi)) {
                sp = bnd4(v[6], v[5], v[4], v[3]);
                sp = dmax(sp, dmin(v[4], v[5]));
                sp = dmin(sp, dmax(v[4], v[5]));
                sm = sedger;
                cw_limit(v[4], sm, sp);
            }
        }
    } else {
        double sedgel = cs_face(v[1], v[2], v[3], v[4]);
        double sedge = cs_face(v[2], v[3], v[4], v[5]);
        double sedger = cs_face(v[3], v[4], v[5], v[6]);
        double sedgerr = cs_face(v[4], v[5], v[6], v[7]);
        cs_cell(v, sedgel, sedge, sedger, sedgerr, 1.e-10, sm, sp);
        if (is_ho(bclo)) {
            if (pos == dlo) {
                if (D == 1) sedge = v[3];
                sm = v[3];
                sp = bnd4(v[3], v[4], v[5], v[6]);
                sp = dmax(sp, dmin(v[5], v[4]));
                sp = dmin(sp, dmax(v[5], v[4]));
            } else if (pos == dlo + 1) {
                sedgel = v[2];
                sedge = bnd4(v[2], v[3], v[4], v[5]);
                sedge = dmax(sedge, dmin(v[4], v[3]));
                sedge = dmin(sedge, dmax(v[4], v[3]));
            } else if (pos == dlo + 2) {
                sedgel = bnd4(v[1], v[2], v[3], v[4]);
                sedgel = dmax(sedgel, dmin(v[3], v[2]));
                sedgel = dmin(sedgel, dmax(v[3], v[2]));
            }
            if (pos == dlo + 1 || pos == dlo + 2) {
                cs_cell(v, sedgel, sedge, sedger, sedgerr, D == 0 ? 1.e10 : 1.e-10, sm, sp);
            }
        }
        if (is_ho(bchi)) {
            if (pos == dhi) {
                sp = v[5];
                sm = bnd4(v[5], v[4], v[3], v[2]);
                sm = dmax(sm, dmin(v[3], v[4]));
                sm = dmin(sm, dmax(v[3], v[4]));
            } else if (pos == dhi - 1) {
                if (D != 0) sedgerr = v[6];
                sedger = bnd4(v[6], v[5], v[4], v[3]);
                sedger = dmax(sedger, dmin(v[4], v[5]));
                sedger = dmin(sedger, dmax(v[4], v[5]));
            } else if (pos == dhi - 2) {
                sedgerr = bnd4(v[7], v[6], v[5], v[4]);
                sedgerr = dmax(sedgerr, dmin(v[5], v[6]));
                sedgerr = dmin(sedgerr, dmax(v[5], v[6]));
            }
            if (pos == dhi - 1 || pos == dhi - 2) {
                cs_cell(v, sedgel, sedge, sedger, sedgerr, 1.e-10, sm, sp);
            }
        }
    }
    sm_out = sm;
    sp_out = sp;
}

// characteristic trace (MaestroPPM.cpp:207-251): Ip uses the velocity at the
// high side, Im at the low side (face velocities) or the cell velocity twice
__device__ __forceinline__ void ppm_trace(double s0, double sm, double sp, double u_hi,
                                          double u_lo, double dt, double dxd, double rel_eps,
                                          double& Ip, double& Im) {
    double s6 = 6.0 * s0 - 3.0 * (sm + sp);
    double sigma = dabs(u_hi) * dt / dxd;
    if (u_hi > rel_eps) {
        Ip = sp - 0.5 * sigma * (sp - sm - (1.0 - 2.0 / 3.0 * sigma) * s6);
    } else {
        Ip = s0;
    }
    sigma = dabs(u_lo) * dt / dxd;
    if (u_lo < -rel_eps) {
        Im = sm + 0.5 * sigma * (sp - sm + (1.0 - 2.0 / 3.0 * sigma) * s6);
    } else {
        Im = s0;
    }
}

// PPM type 1 face value between cells sa (left) and sb (right) from the two
// cells' van Leer slopes (MaestroPPM.cpp:86-99: sm of the right cell, which
// equals sp of the left cell)
__device__ __forceinline__ double face1(double sa, double sb, double dla, double dlb) {
    double e = 0.5 * (sb + sa) - (dlb - dla) / 6.0;
    e = dmax(e, dmin(sb, sa));
    e = dmin(e, dmax(sb, sa));
    return e;
}

// does the cell at `pos` take a PPM boundary stencil?
template <int PPM>
__device__ __forceinline__ bool ppm_special(int pos, int dlo, int dhi, int bclo, int bchi) {
    constexpr int R = PPM == 2 ? 2 : 1;
    return (is_ho(bclo) && pos >= dlo && pos <= dlo + R) ||
           (is_ho(bchi) && pos <= dhi && pos >= dhi - R);
}

// the characteristic trace of ppm_trace with sigma of each face given
__device__ __forceinline__ void trace_sig(double s0, double sm, double sp, double u_hi,
                                          double u_lo, double sig_hi, double sig_lo,
                                          double rel_eps, double& Ip, double& Im) {
    double s6 = 6.0 * s0 - 3.0 * (sm + sp);
    if (u_hi > rel_eps) {
        Ip = sp - 0.5 * sig_hi * (sp - sm - (1.0 - 2.0 / 3.0 * sig_hi) * s6);
    } else {
        Ip = s0;
    }
    if (u_lo < -rel_eps) {
        Im = sm + 0.5 * sig_lo * (sp - sm + (1.0 - 2.0 / 3.0 * sig_lo) * s6);
    } else {
        Im = s0;
    }
}

// load the 9-point stencil along a direction
template <int PPM>
__device__ __forceinline__ void load_stencil(const double* q, long long st, double* v) {
    if (PPM == 2) {
        v[1] = q[-3 * st];
        v[7] = q[3 * st];
    } else {
        v[1] = 0.0;
        v[7] = 0.0;
    }
    v[0] = 0.0;
    v[8] = 0.0;
    v[2] = q[-2 * st];
    v[3] = q[-st];
    v[4] = q[0];
    v[5] = q[st];
    v[6] = q[2 * st];
}

// ---------------------------------------------------------------------------
// limited slopes (MaestroSlopes.cpp); v[o+4] = s(cell+o), needs v[2..6]
template <int D>
__device__ __forceinline__ double slope_point(const double* v, int order, int pos, int ilo,
                                              int ihi, int bclo, int bchi) {
    double out = 0.0;
    const bool lo_ho = is_ho(bclo), hi_ho = is_ho(bchi);
    if (order == 0) {
        out = 0.0;
    } else if (order == 2) {
        double del = 0.5 * (v[5] - v[3]);
        double dpls = 2.0 * (v[5] - v[4]);
        double dmn = 2.0 * (v[4] - v[3]);
        double slim = nn_min(dabs(dpls), dabs(dmn));
        slim = dpls * dmn > 0.0 ? slim : 0.0;
        double sflag = dsign1(del);
        out = sflag * dmin(slim, dabs(del));
        if (lo_ho) {
            if (pos == ilo - 1) {
                out = 0.0;
            } else if (pos == ilo) {
                del = (v[5] + 3.0 * v[4] - 4.0 * v[3]) / 3.0;
                dpls = 2.0 * (v[5] - v[4]);
                dmn = 2.0 * (v[4] - v[3]);
                slim = nn_min(dabs(dpls), dabs(dmn));
                slim = dpls * dmn > 0.0 ? slim : 0.0;
                sflag = dsign1(del);
                out = sflag * dmin(slim, dabs(del));
            }
        }
        if (hi_ho) {
            if (pos == ihi + 1) {
                out = 0.0;
            } else if (pos == ihi) {
                del = -(v[3] + 3.0 * v[4] - 4.0 * v[5]) / 3.0;
                if (D == 0) {
                    dpls = 2.0 * (v[4] - v[3]);
                    dmn = 2.0 * (v[5] - v[4]);
                } else {
                    dpls = 2.0 * (v[5] - v[4]);
                    dmn = 2.0 * (v[4] - v[3]);
                }
                slim = nn_min(dabs(dpls), dabs(dmn));
                slim = dpls * dmn > 0.0 ? slim : 0.0;
                sflag = dsign1(del);
                out = sflag * dmin(slim, dabs(del));
            }
        }
    } else if (order == 4) {
        double dcen = 0.5 * (v[4] - v[2]);
        double dmn = 2.0 * (v[3] - v[2]);
        double dpls = 2.0 * (v[4] - v[3]);
        double dlim = nn_min(dabs(dmn), dabs(dpls));
        dlim = dpls * dmn > 0.0 ? dlim : 0.0;
        double dflag = dsign1(dcen);
        double dxl = dflag * dmin(dlim, dabs(dcen));
        dcen = 0.5 * (v[6] - v[4]);
        dmn = 2.0 * (v[5] - v[4]);
        dpls = 2.0 * (v[6] - v[5]);
        dlim = nn_min(dabs(dmn), dabs(dpls));
        dlim = dpls * dmn > 0.0 ? dlim : 0.0;
        dflag = dsign1(dcen);
        double dxr = dflag * dmin(dlim, dabs(dcen));
        dcen = 0.5 * (v[5] - v[3]);
        dmn = 2.0 * (v[4] - v[3]);
        dpls = 2.0 * (v[5] - v[4]);
        dlim = nn_min(dabs(dmn), dabs(dpls));
        dlim = dpls * dmn > 0.0 ? dlim : 0.0;
        dflag = dsign1(dcen);
        double ds = 4.0 / 3.0 * dcen - (dxr + dxl) / 6.0;
        out = dflag * nn_min(dabs(ds), dlim);
        if (lo_ho) {
            if (pos == ilo - 1) {
                out = 0.0;
            } else if (pos == ilo) {
                double del = -16.0 / 15.0 * v[3] + 0.5 * v[4] + 2.0 / 3.0 * v[5] - 0.1 * v[6];
                dmn = 2.0 * (v[4] - v[3]);
                dpls = 2.0 * (v[5] - v[4]);
                double slim = nn_min(dabs(dpls), dabs(dmn));
                slim = dpls * dmn > 0.0 ? (D == 0 ? dlim : slim) : 0.0;
                double sflag = dsign1(del);
                out = sflag * dmin(slim, dabs(del));
            } else if (pos == ilo + 1) {
                double del = -16.0 / 15.0 * v[2] + 0.5 * v[3] + 2.0 / 3.0 * v[4] - 0.1 * v[5];
                dmn = 2.0 * (v[3] - v[2]);
                dpls = 2.0 * (v[4] - v[3]);
                double slim = nn_min(dabs(dpls), dabs(dmn));
                slim = dpls * dmn > 0.0 ? (D == 0 ? dlim : slim) : 0.0;
                double sflag = dsign1(del);
                dxl = sflag * dmin(slim, dabs(del));
                ds = 4.0 / 3.0 * dcen - (dxr + dxl) / 6.0;
                out = dflag * nn_min(dabs(ds), dlim);
            }
        }
        if (hi_ho) {
            if (pos == ihi + 1) {
                out = 0.0;
            } else if (pos == ihi) {
                double del = -(-16.0 / 15.0 * v[5] + 0.5 * v[4] + 2.0 / 3.0 * v[3] - 0.1 * v[2]);
                dmn = 2.0 * (v[4] - v[3]);
                dpls = 2.0 * (v[5] - v[4]);
                double slim = nn_min(dabs(dpls), dabs(dmn));
                slim = dpls * dmn > 0.0 ? (D == 0 ? dlim : slim) : 0.0;
                double sflag = dsign1(del);
                out = sflag * dmin(slim, dabs(del));
            } else if (pos == ihi - 1) {
                double del = -(-16.0 / 15.0 * v[6] + 0.5 * v[5] + 2.0 / 3.0 * v[4] - 0.1 * v[3]);
                dmn = 2.0 * (v[5] - v[4]);
                dpls = 2.0 * (v[6] - v[5]);
                double slim = nn_min(dabs(dpls), dabs(dmn));
                slim = dpls * dmn > 0.0 ? (D == 0 ? dlim : slim) : 0.0;
                double sflag = dsign1(del);
                dxr = sflag * dmin(slim, dabs(del));
                ds = 4.0 / 3.0 * dcen - (dxl + dxr) / 6.0;
                out = dflag * nn_min(dabs(ds), dlim);
            }
        }
    }
    return out;
}

// ---------------------------------------------------------------------------
// kernel parameters shared by every stage
struct KGeom {
    int dim;
    int ppm_type;
    int slope_order;
    int ppm_trace_forces;
    int spherical;
    double dt;
    double rel_eps;
    double dx[3];
    int domlo[3];
    int domhi[3];
    int phys_lo[3];
    int phys_hi[3];
    int xcd;  // 1: grids from grid_for_g (x a multiple of 8), indices from grid_point
};

// The linear point index of this thread in a grid_for_g launch.  Workgroups
// are dispatched to the 8 XCDs round robin by linear id; with the grid's x
// extent a multiple of 8 the XCD is blockIdx.x % 8, and the remap gives each
// XCD one contiguous eighth of a box (its stencil planes stay in that XCD's
// L2) instead of every eighth block.
__device__ __forceinline__ long long grid_point(const KGeom& g) {
    unsigned b = blockIdx.x;
    if (g.xcd) b = (b & 7u) * (gridDim.x >> 3) + (b >> 3);
    return (long long)b * blockDim.x + threadIdx.x;
}

// XCD-aware orders of a per-component tile launch over a 1-D grid of
// 8 * nc * ceil(ntile / 8) blocks (xcd_tile_blocks).  XCD x =
// blockIdx.x % 8 runs the nc component blocks of each of its tiles back to
// back, so the planes they all stage (the velocities) are read from HBM
// once into that XCD's L2.  Its tiles are
//   mode 1: x, x + 8, x + 16, ... (the round robin of the launch order,
//           which balances tiles of unequal cost);
//   mode 2: the contiguous range [x * per, (x + 1) * per), per =
//           ceil(ntile / 8) (neighbouring tiles share their halo in L2).
// False for the blocks past the last tile.
__device__ __forceinline__ bool xcd_tile(int ntile, int nc, int mode, int& t, int& v) {
    const unsigned L = blockIdx.x;
    const unsigned r = L >> 3;
    v = (int)(r % (unsigned)nc);
    const unsigned local = r / (unsigned)nc;
    if (mode == 2)
        t = (int)((L & 7u) * (((unsigned)ntile + 7u) / 8u) + local);
    else
        t = (int)(local * 8u + (L & 7u));
    return t < ntile;
}
__host__ __device__ constexpr unsigned xcd_tile_blocks(long long ntile, int nc) {
    return (unsigned)(8LL * nc * ((ntile + 7) / 8));
}

// region iteration: linear index -> (i, j, k) in box [lo, hi]
__device__ __forceinline__ bool region_point(long long t, const int* lo, const int* hi, int& i,
                                             int& j, int& k) {
    const unsigned nx = (unsigned)(hi[0] - lo[0] + 1);
    const unsigned ny = (unsigned)(hi[1] - lo[1] + 1);
    const unsigned nz = (unsigned)(hi[2] - lo[2] + 1);
    if (t >= (long long)nx * ny * nz) return false;
    const unsigned u = (unsigned)t;
    const unsigned kk = u / (nx * ny);
    const unsigned r = u - kk * nx * ny;
    const unsigned jj = r / nx;
    i = lo[0] + (int)(r - jj * nx);
    j = lo[1] + (int)jj;
    k = lo[2] + (int)kk;
    return true;
}

}  // namespace mxd

#endif
