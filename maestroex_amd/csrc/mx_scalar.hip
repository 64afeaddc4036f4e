// mx_scalar.hip -- MakeEdgeScal, MakeRhoXFlux, MakeRhoHFlux, BCSetup and the
// whole advection pass on the GPU.
//
// MakeEdgeScal (MaestroMakeEdgeScalars.cpp:6-1667) runs as three kernels over
// every box and every requested component at once:
//   k_es_ppm     PPM (or slopes) of the scalar in every direction on the box
//                grown by one, traced with the MAC velocities (Ip/Im);
//   k_es_trans   the six corner-coupled states simhxy..simhzy (3-D), with the
//                predictor states slx..srz and simhx..simhz evaluated on the
//                fly from Ip/Im, the MAC velocities and the boundary types;
//   k_es_edge    the edge states on the valid faces (sedge).
// The fluxes are one kernel each over the valid faces of all boxes.
#include <algorithm>

#include "mx_internal.H"

using namespace mxd;
using namespace mxh;

namespace {

constexpr int e(int d, int q) { return d == q ? 1 : 0; }
__host__ __device__ constexpr int pair_index(int d, int t) { return d * 2 + (t > d ? t - 1 : t); }

constexpr int MAXC = MX_MAX_SCAL;

struct EsBox {
    int lo[3], hi[3];
    DA s, fo;
    DA um[3], se[3];
};

struct EsTmp {
    DA Ip, Im, Ipf, Imf;  // ppm_type 0: Ip holds the slope along each direction
    DA T[6];
    int bclo[3], bchi[3];  // bcs[bccomp] of this component
};

struct EsArgs {
    int start_scomp;
    int num_comp;
    int is_vel;
    int is_conservative;
};

template <int PPM, int D>
__device__ __forceinline__ void es_ppm_dir(const EsBox& B, const EsTmp& Tm, const KGeom& g,
                                           const EsArgs& A, int vc, int i, int j, int k) {
    const int comp = A.start_scomp + vc;
    const int pos = D == 0 ? i : (D == 1 ? j : k);
    const int bclo = Tm.bclo[D], bchi = Tm.bchi[D];
    double v[9];
    if (PPM == 0) {
        load_stencil<1>(B.s.p + B.s.off(i, j, k, comp), B.s.stride(D), v);
        Tm.Ip(i, j, k, D) = slope_point<D>(v, g.slope_order, pos, g.domlo[D], g.domhi[D], bclo, bchi);
        return;
    }
    double sm, sp, Ip, Im;
    const double uhi = B.um[D](i + e(D, 0), j + e(D, 1), k + e(D, 2));
    const double ulo = B.um[D](i, j, k);
    load_stencil<PPM>(B.s.p + B.s.off(i, j, k, comp), B.s.stride(D), v);
    ppm_edges<PPM, D>(v, pos, g.domlo[D], g.domhi[D], bclo, bchi, sm, sp);
    ppm_trace(v[4], sm, sp, uhi, ulo, g.dt, g.dx[D], g.rel_eps, Ip, Im);
    Tm.Ip(i, j, k, D) = Ip;
    Tm.Im(i, j, k, D) = Im;
    if (g.ppm_trace_forces == 1) {
        load_stencil<PPM>(B.fo.p + B.fo.off(i, j, k, comp), B.fo.stride(D), v);
        ppm_edges<PPM, D>(v, pos, g.domlo[D], g.domhi[D], bclo, bchi, sm, sp);
        ppm_trace(v[4], sm, sp, uhi, ulo, g.dt, g.dx[D], g.rel_eps, Ip, Im);
        Tm.Ipf(i, j, k, D) = Ip;
        Tm.Imf(i, j, k, D) = Im;
    }
}

template <int PPM>
__global__ void __launch_bounds__(256) k_es_ppm(const EsBox* __restrict__ boxes,
                                                const EsTmp* __restrict__ tmps, KGeom g, EsArgs A) {
    const EsBox& B = boxes[blockIdx.y];
    const int vc = blockIdx.z;
    const EsTmp& Tm = tmps[blockIdx.y * A.num_comp + vc];
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = B.lo[d] - (d < g.dim ? 1 : 0);
        hi[d] = B.hi[d] + (d < g.dim ? 1 : 0);
    }
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!region_point(idx, lo, hi, i, j, k)) return;
    es_ppm_dir<PPM, 0>(B, Tm, g, A, vc, i, j, k);
    es_ppm_dir<PPM, 1>(B, Tm, g, A, vc, i, j, k);
    if (g.dim == 3) es_ppm_dir<PPM, 2>(B, Tm, g, A, vc, i, j, k);
}

// predictor states sl/sr on D-faces (MakeEdgeScalPredictor)
template <int D>
__device__ __forceinline__ void es_pred(const EsBox& B, const EsTmp& Tm, const KGeom& g,
                                        const EsArgs& A, int vc, int i, int j, int k, double& l,
                                        double& r) {
    const int comp = A.start_scomp + vc;
    const int im = i - e(D, 0), jm = j - e(D, 1), km = k - e(D, 2);
    if (g.ppm_type == 0) {
        const double u = B.um[D](i, j, k);
        const double h = g.dx[D];
        if (g.dim == 3) {
            l = B.s(im, jm, km, comp) + 0.5 * (1.0 - g.dt * u / h) * Tm.Ip(im, jm, km, D);
            r = B.s(i, j, k, comp) - 0.5 * (1.0 + g.dt * u / h) * Tm.Ip(i, j, k, D);
        } else {
            const double dt2 = 0.5 * g.dt;
            l = B.s(im, jm, km, comp) + (0.5 - dt2 * u / h) * Tm.Ip(im, jm, km, D);
            r = B.s(i, j, k, comp) - (0.5 + dt2 * u / h) * Tm.Ip(i, j, k, D);
        }
    } else {
        l = Tm.Ip(im, jm, km, D);
        r = Tm.Im(i, j, k, D);
    }
    const int pos = D == 0 ? i : (D == 1 ? j : k);
    if (pos == g.domlo[D]) {
        const int bc = Tm.bclo[D];
        if (bc == MX_BC_EXT_DIR) {
            l = B.s(im, jm, km, comp);
            r = l;
        } else if (bc == MX_BC_FOEXTRAP || bc == MX_BC_HOEXTRAP) {
            if (A.is_vel && comp == D) r = dmin(r, 0.0);
            l = r;
        } else if (bc == MX_BC_REFLECT_EVEN) {
            l = r;
        } else if (bc == MX_BC_REFLECT_ODD) {
            l = 0.0;
            r = 0.0;
        }
    } else if (pos == g.domhi[D] + 1) {
        const int bc = Tm.bchi[D];
        if (bc == MX_BC_EXT_DIR) {
            l = B.s(i, j, k, comp);
            r = l;
        } else if (bc == MX_BC_FOEXTRAP || bc == MX_BC_HOEXTRAP) {
            if (A.is_vel && comp == D) l = dmax(l, 0.0);
            r = l;
        } else if (bc == MX_BC_REFLECT_EVEN) {
            r = l;
        } else if (bc == MX_BC_REFLECT_ODD) {
            l = 0.0;
            r = 0.0;
        }
    }
}

template <int D>
__device__ __forceinline__ double es_simh(const EsBox& B, const EsTmp& Tm, const KGeom& g,
                                          const EsArgs& A, int vc, int i, int j, int k) {
    double l, r;
    es_pred<D>(B, Tm, g, A, vc, i, j, k, l, r);
    const double u = B.um[D](i, j, k);
    const double thr = g.dim == 3 ? 0.0 : g.rel_eps;
    double res = (u > 0.0) ? l : r;
    res = (dabs(u) > thr) ? res : 0.5 * (l + r);
    return res;
}

__device__ __forceinline__ double es_divu(const EsBox& B, const KGeom& g, int i, int j, int k) {
    double dv = B.um[0](i + 1, j, k) - B.um[0](i, j, k) + B.um[1](i, j + 1, k) -
                B.um[1](i, j, k) + B.um[2](i, j, k + 1) - B.um[2](i, j, k);
    dv /= g.dx[0];
    return dv;
}

// corner-coupled state on D-faces corrected by T (MakeEdgeScalTransverse)
template <int D, int T>
__device__ __forceinline__ void es_trans(const EsBox& B, const EsTmp& Tm, const KGeom& g,
                                         const EsArgs& A, int vc, int i, int j, int k) {
    const int comp = A.start_scomp + vc;
    const int im = i - e(D, 0), jm = j - e(D, 1), km = k - e(D, 2);
    const int it = e(T, 0), jt = e(T, 1), kt = e(T, 2);
    const double ht = g.dx[T];
    const DA& ut = B.um[T];
    double l, r;
    es_pred<D>(B, Tm, g, A, vc, i, j, k, l, r);
    if (A.is_conservative) {
        const double dt3 = g.dt / 3.0;
        l = l -
            (dt3 / ht) * (es_simh<T>(B, Tm, g, A, vc, im + it, jm + jt, km + kt) *
                              ut(im + it, jm + jt, km + kt) -
                          es_simh<T>(B, Tm, g, A, vc, im, jm, km) * ut(im, jm, km)) -
            dt3 * B.s(im, jm, km, comp) * es_divu(B, g, im, jm, km) +
            (dt3 / ht) * B.s(im, jm, km, comp) * (ut(im + it, jm + jt, km + kt) - ut(im, jm, km));
        r = r -
            (dt3 / ht) * (es_simh<T>(B, Tm, g, A, vc, i + it, j + jt, k + kt) *
                              ut(i + it, j + jt, k + kt) -
                          es_simh<T>(B, Tm, g, A, vc, i, j, k) * ut(i, j, k)) -
            dt3 * B.s(i, j, k, comp) * es_divu(B, g, i, j, k) +
            (dt3 / ht) * B.s(i, j, k, comp) * (ut(i + it, j + jt, k + kt) - ut(i, j, k));
    } else {
        const double dt6 = g.dt / 6.0;
        l = l - (dt6 / ht) * (ut(im + it, jm + jt, km + kt) + ut(im, jm, km)) *
                    (es_simh<T>(B, Tm, g, A, vc, im + it, jm + jt, km + kt) -
                     es_simh<T>(B, Tm, g, A, vc, im, jm, km));
        r = r - (dt6 / ht) * (ut(i + it, j + jt, k + kt) + ut(i, j, k)) *
                    (es_simh<T>(B, Tm, g, A, vc, i + it, j + jt, k + kt) -
                     es_simh<T>(B, Tm, g, A, vc, i, j, k));
    }
    const int pos = D == 0 ? i : (D == 1 ? j : k);
    if (pos == g.domlo[D]) {
        const int bc = Tm.bclo[D];
        if (bc == MX_BC_EXT_DIR) {
            l = B.s(im, jm, km, comp);
            r = l;
        } else if (bc == MX_BC_FOEXTRAP || bc == MX_BC_HOEXTRAP) {
            if (A.is_vel && comp == D) r = dmin(r, 0.0);
            l = r;
        } else if (bc == MX_BC_REFLECT_EVEN) {
            l = r;
        } else if (bc == MX_BC_REFLECT_ODD) {
            l = 0.0;
            r = 0.0;
        }
    } else if (pos == g.domhi[D] + 1) {
        const int bc = Tm.bchi[D];
        if (bc == MX_BC_EXT_DIR) {
            l = B.s(i, j, k, comp);
            r = l;
        } else if (bc == MX_BC_FOEXTRAP || bc == MX_BC_HOEXTRAP) {
            if (A.is_vel && comp == D) l = dmax(l, 0.0);
            r = l;
        } else if (bc == MX_BC_REFLECT_EVEN) {
            r = l;
        } else if (bc == MX_BC_REFLECT_ODD) {
            l = 0.0;
            r = 0.0;
        }
    }
    const double u = B.um[D](i, j, k);
    double res = (u > 0.0) ? l : r;
    res = (dabs(u) > g.rel_eps) ? res : 0.5 * (l + r);
    Tm.T[pair_index(D, T)](i, j, k) = res;
}

__global__ void __launch_bounds__(256) k_es_trans(const EsBox* __restrict__ boxes,
                                                  const EsTmp* __restrict__ tmps, KGeom g,
                                                  EsArgs A) {
    const EsBox& B = boxes[blockIdx.y];
    const int vc = blockIdx.z / 6;
    const int p = blockIdx.z % 6;
    const EsTmp& Tm = tmps[blockIdx.y * A.num_comp + vc];
    const int D = p / 2;
    const int q = p % 2;
    const int T = (q < D) ? q : q + 1;
    const int O = 3 - D - T;
    int lo[3] = {B.lo[0], B.lo[1], B.lo[2]};
    int hi[3] = {B.hi[0], B.hi[1], B.hi[2]};
    lo[O] -= 1;
    hi[O] += 1;
    hi[D] += 1;
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!region_point(idx, lo, hi, i, j, k)) return;
    switch (p) {
        case 0: es_trans<0, 1>(B, Tm, g, A, vc, i, j, k); break;
        case 1: es_trans<0, 2>(B, Tm, g, A, vc, i, j, k); break;
        case 2: es_trans<1, 0>(B, Tm, g, A, vc, i, j, k); break;
        case 3: es_trans<1, 2>(B, Tm, g, A, vc, i, j, k); break;
        case 4: es_trans<2, 0>(B, Tm, g, A, vc, i, j, k); break;
        default: es_trans<2, 1>(B, Tm, g, A, vc, i, j, k); break;
    }
}

// one transverse term of MakeEdgeScalEdges for direction T
template <int D, int T>
__device__ __forceinline__ void es_edge_term(const EsBox& B, const EsTmp& Tm, const KGeom& g,
                                             const EsArgs& A, int vc, int i, int j, int k,
                                             double& l, double& r) {
    const int im = i - e(D, 0), jm = j - e(D, 1), km = k - e(D, 2);
    const int it = e(T, 0), jt = e(T, 1), kt = e(T, 2);
    const double ht = g.dx[T];
    const DA& ut = B.um[T];
    double aL, bL, aR, bR;
    if (g.dim == 3) {
        const DA& Tt = Tm.T[pair_index(T, 3 - D - T)];
        aL = Tt(im + it, jm + jt, km + kt);
        bL = Tt(im, jm, km);
        aR = Tt(i + it, j + jt, k + kt);
        bR = Tt(i, j, k);
    } else {
        aL = es_simh<T>(B, Tm, g, A, vc, im + it, jm + jt, km + kt);
        bL = es_simh<T>(B, Tm, g, A, vc, im, jm, km);
        aR = es_simh<T>(B, Tm, g, A, vc, i + it, j + jt, k + kt);
        bR = es_simh<T>(B, Tm, g, A, vc, i, j, k);
    }
    if (A.is_conservative) {
        const double dt2 = 0.5 * g.dt;
        l = l - (dt2 / ht) * (aL * ut(im + it, jm + jt, km + kt) - bL * ut(im, jm, km));
        r = r - (dt2 / ht) * (aR * ut(i + it, j + jt, k + kt) - bR * ut(i, j, k));
    } else {
        const double dt4 = 0.25 * g.dt;
        l = l - (dt4 / ht) * (ut(im + it, jm + jt, km + kt) + ut(im, jm, km)) * (aL - bL);
        r = r - (dt4 / ht) * (ut(i + it, j + jt, k + kt) + ut(i, j, k)) * (aR - bR);
    }
}

template <int D>
__device__ __forceinline__ void es_edge(const EsBox& B, const EsTmp& Tm, const KGeom& g,
                                        const EsArgs& A, int vc, int i, int j, int k) {
    const int comp = A.start_scomp + vc;
    const int im = i - e(D, 0), jm = j - e(D, 1), km = k - e(D, 2);
    const double dt2 = 0.5 * g.dt;
    double fl, fr;
    if (g.ppm_trace_forces == 0) {
        fl = B.fo(im, jm, km, comp);
        fr = B.fo(i, j, k, comp);
    } else {
        fl = Tm.Ipf(im, jm, km, D);
        fr = Tm.Imf(i, j, k, D);
    }
    double l, r;
    es_pred<D>(B, Tm, g, A, vc, i, j, k, l, r);
    if (D == 0) {
        es_edge_term<0, 1>(B, Tm, g, A, vc, i, j, k, l, r);
        if (g.dim == 3) es_edge_term<0, 2>(B, Tm, g, A, vc, i, j, k, l, r);
    } else if (D == 1) {
        es_edge_term<1, 0>(B, Tm, g, A, vc, i, j, k, l, r);
        if (g.dim == 3) es_edge_term<1, 2>(B, Tm, g, A, vc, i, j, k, l, r);
    } else {
        es_edge_term<2, 0>(B, Tm, g, A, vc, i, j, k, l, r);
        es_edge_term<2, 1>(B, Tm, g, A, vc, i, j, k, l, r);
    }
    const DA& ud = B.um[D];
    if (A.is_conservative) {
        const double hd = g.dx[D];
        l = l - (dt2 / hd) * B.s(im, jm, km, comp) * (ud(i, j, k) - ud(im, jm, km));
        r = r - (dt2 / hd) * B.s(i, j, k, comp) *
                    (ud(i + e(D, 0), j + e(D, 1), k + e(D, 2)) - ud(i, j, k));
    }
    const double sedgel = l + dt2 * fl;
    const double sedger = r + dt2 * fr;
    const double u = ud(i, j, k);
    double res = (u > 0.0) ? sedgel : sedger;
    res = (dabs(u) > g.rel_eps) ? res : 0.5 * (sedgel + sedger);
    const int pos = D == 0 ? i : (D == 1 ? j : k);
    if (pos == g.domlo[D]) {
        const int bc = Tm.bclo[D];
        if (bc == MX_BC_EXT_DIR) {
            res = B.s(im, jm, km, comp);
        } else if (bc == MX_BC_FOEXTRAP || bc == MX_BC_HOEXTRAP) {
            res = (A.is_vel && comp == D) ? dmin(sedger, 0.0) : sedger;
        } else if (bc == MX_BC_REFLECT_EVEN) {
            res = sedger;
        } else if (bc == MX_BC_REFLECT_ODD) {
            res = 0.0;
        }
    } else if (pos == g.domhi[D] + 1) {
        const int bc = Tm.bchi[D];
        if (bc == MX_BC_EXT_DIR) {
            res = B.s(i, j, k, comp);
        } else if (bc == MX_BC_FOEXTRAP || bc == MX_BC_HOEXTRAP) {
            res = (A.is_vel && comp == D) ? dmax(sedgel, 0.0) : sedgel;
        } else if (bc == MX_BC_REFLECT_EVEN) {
            res = sedgel;
        } else if (bc == MX_BC_REFLECT_ODD) {
            res = 0.0;
        }
    }
    B.se[D](i, j, k, comp) = res;
}

__global__ void __launch_bounds__(256) k_es_edge(const EsBox* __restrict__ boxes,
                                                 const EsTmp* __restrict__ tmps, KGeom g,
                                                 EsArgs A) {
    const EsBox& B = boxes[blockIdx.y];
    const int vc = blockIdx.z / g.dim;
    const int d = blockIdx.z % g.dim;
    const EsTmp& Tm = tmps[blockIdx.y * A.num_comp + vc];
    int lo[3] = {B.lo[0], B.lo[1], B.lo[2]};
    int hi[3] = {B.hi[0], B.hi[1], B.hi[2]};
    hi[d] += 1;
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!region_point(idx, lo, hi, i, j, k)) return;
    if (d == 0)
        es_edge<0>(B, Tm, g, A, vc, i, j, k);
    else if (d == 1)
        es_edge<1>(B, Tm, g, A, vc, i, j, k);
    else
        es_edge<2>(B, Tm, g, A, vc, i, j, k);
}

// ---------------------------------------------------------------------------
// fluxes

struct FluxBox {
    int lo[3], hi[3];
    DA sf[3], se[3], um[3];
    DA eta;
    DA r0o[3], r0n[3], rh0o[3], rh0n[3], h0o[3], h0n[3];
};

struct BaseP {
    const double* p;
    long long off;  // lev * len
    __device__ __forceinline__ double operator()(int r) const { return p[off + r]; }
};

struct FluxArgs {
    BaseP rho0_old, rho0_edge_old, rho0_new, rho0_edge_new, rho0_pe, w0;
    BaseP rhoh0_old, rhoh0_edge_old, rhoh0_new, rhoh0_edge_new;
    int Rho, RhoH, FirstSpec, NumSpec;
    int start_comp, num_comp;
    int species_pred_type, enthalpy_pred_type;
    int spherical, use_exact, evolve;
};

__global__ void k_zero_fab(const FluxBox* __restrict__ boxes, int nb) {
    (void)nb;
    const FluxBox& B = boxes[blockIdx.y];
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!region_point(idx, B.lo, B.hi, i, j, k)) return;
    B.eta(i, j, k) = 0.0;
}

template <int D>
__device__ __forceinline__ void rhox_face(const FluxBox& B, const KGeom& g, const FluxArgs& A,
                                          int i, int j, int k) {
    const int r = g.dim == 2 ? j : k;
    const bool vertical = (D == g.dim - 1);
    const bool sph = g.dim == 3 && A.spherical;
    const DA& sf = B.sf[D];
    const DA& se = B.se[D];
    const double um = B.um[D](i, j, k);
    // F_Rho and etarhoflux accumulate in registers in the reference's order
    // (0 + F_1 + F_2 + ..., MaestroMakeFlux.cpp:139) and are stored once
    const bool eta_on = vertical && !sph && A.evolve && !A.use_exact;
    double eta = eta_on ? B.eta(i, j, k) : 0.0;
    double frho = 0.0;
    double rho0_edge;
    if (sph) {
        rho0_edge = 0.5 * B.r0o[D](i, j, k) + 0.5 * B.r0n[D](i, j, k);
    } else if (vertical) {
        rho0_edge = 0.5 * (A.rho0_edge_old(r) + A.rho0_edge_new(r));
    } else {
        rho0_edge = 0.5 * (A.rho0_old(r) + A.rho0_new(r));
    }
    const double se_rho = se(i, j, k, A.Rho);
    for (int comp = A.start_comp; comp < A.num_comp + A.start_comp; ++comp) {
        double f;
        if (A.species_pred_type == 1) {
            f = um * (rho0_edge + se_rho) * se(i, j, k, comp);
        } else if (A.species_pred_type == 2) {
            f = um * se(i, j, k, comp);
        } else if (A.species_pred_type == 3) {
            f = um * se_rho * se(i, j, k, comp);
        } else {
            f = sf(i, j, k, comp);  // no flux form: the value is kept
        }
        sf(i, j, k, comp) = f;
        if (eta_on) {
            if (comp >= A.FirstSpec && comp <= A.FirstSpec + A.NumSpec - 1) eta += f;
            if (comp == A.FirstSpec + A.NumSpec - 1) eta -= A.w0(r) * A.rho0_pe(r);
        }
        if (comp == A.Rho) frho = f;  // the component store would have replaced it
        frho += f;
    }
    sf(i, j, k, A.Rho) = frho;
    if (eta_on) B.eta(i, j, k) = eta;
}

__global__ void __launch_bounds__(256) k_rhox_flux(const FluxBox* __restrict__ boxes, KGeom g,
                                                   FluxArgs A) {
    const FluxBox& B = boxes[blockIdx.y];
    const int d = blockIdx.z;
    int lo[3] = {B.lo[0], B.lo[1], B.lo[2]};
    int hi[3] = {B.hi[0], B.hi[1], B.hi[2]};
    hi[d] += 1;
    long long idx = grid_point(g);
    int i, j, k;
    if (!region_point(idx, lo, hi, i, j, k)) return;
    if (d == 0)
        rhox_face<0>(B, g, A, i, j, k);
    else if (d == 1)
        rhox_face<1>(B, g, A, i, j, k);
    else
        rhox_face<2>(B, g, A, i, j, k);
}

template <int D>
__device__ __forceinline__ void rhoh_face(const FluxBox& B, const KGeom& g, const FluxArgs& A,
                                          int i, int j, int k) {
    const int r = g.dim == 2 ? j : k;
    const bool vertical = (D == g.dim - 1);
    const bool sph = g.dim == 3 && A.spherical;
    const int ept = A.enthalpy_pred_type;
    const int spt = A.species_pred_type;
    const bool have_h = ept == 2 || ept == 4 || ept == 6;
    const bool have_hprime = ept == 5;
    const bool have_rhoh = ept == 0;
    const DA& sf = B.sf[D];
    const DA& se = B.se[D];
    const double um = B.um[D](i, j, k);
    const int Rho = A.Rho, RhoH = A.RhoH;
    if (!sph) {
        if (have_h) {
            if (spt == 1) {
                double rho0_edge = vertical ? 0.5 * (A.rho0_edge_old(r) + A.rho0_edge_new(r))
                                            : 0.5 * (A.rho0_old(r) + A.rho0_new(r));
                sf(i, j, k, RhoH) = um * (rho0_edge + se(i, j, k, Rho)) * se(i, j, k, RhoH);
            } else if (spt == 3 || spt == 2) {
                sf(i, j, k, RhoH) = um * se(i, j, k, Rho) * se(i, j, k, RhoH);
            }
        } else if (have_rhoh) {
            sf(i, j, k, RhoH) = um * se(i, j, k, RhoH);
        } else if (ept == 1 || ept == 3) {
            double rhoh0_edge = vertical ? 0.5 * (A.rhoh0_edge_old(r) + A.rhoh0_edge_new(r))
                                         : 0.5 * (A.rhoh0_old(r) + A.rhoh0_new(r));
            sf(i, j, k, RhoH) = um * (rhoh0_edge + se(i, j, k, RhoH));
        }
    } else if (A.use_exact) {
        double rhoh0_edge = 0.5 * B.rh0o[D](i, j, k) + 0.5 * B.rh0n[D](i, j, k);
        if (D == 0) {
            if (!(have_h || have_rhoh)) sf(i, j, k, RhoH) = um * (rhoh0_edge + se(i, j, k, RhoH));
        } else {
            if (have_h || have_hprime) {
            } else if (have_rhoh) {
                sf(i, j, k, RhoH) = um * se(i, j, k, RhoH);
            } else {
                sf(i, j, k, RhoH) = um * (rhoh0_edge + se(i, j, k, RhoH));
            }
        }
    } else {
        double rho0_edge = 0.5 * B.r0o[D](i, j, k) + 0.5 * B.r0n[D](i, j, k);
        double h0_edge = 0.5 * B.h0o[D](i, j, k) + 0.5 * B.h0n[D](i, j, k);
        if (have_h) {
            if (spt == 1) {
                sf(i, j, k, RhoH) = um * (rho0_edge + se(i, j, k, Rho)) * se(i, j, k, RhoH);
            } else if (spt == 3 || spt == 2) {
                sf(i, j, k, RhoH) = um * se(i, j, k, Rho) * se(i, j, k, RhoH);
            }
        } else if (have_hprime) {
            if (spt == 1) {
                sf(i, j, k, RhoH) =
                    um * (se(i, j, k, Rho) + rho0_edge) * (se(i, j, k, RhoH) + h0_edge);
            }
        } else if (have_rhoh) {
            sf(i, j, k, RhoH) = um * se(i, j, k, RhoH);
        } else {
            sf(i, j, k, RhoH) = um * (rho0_edge * h0_edge + se(i, j, k, RhoH));
        }
    }
}

__global__ void __launch_bounds__(256) k_rhoh_flux(const FluxBox* __restrict__ boxes, KGeom g,
                                                   FluxArgs A) {
    const FluxBox& B = boxes[blockIdx.y];
    const int d = blockIdx.z;
    int lo[3] = {B.lo[0], B.lo[1], B.lo[2]};
    int hi[3] = {B.hi[0], B.hi[1], B.hi[2]};
    hi[d] += 1;
    long long idx = grid_point(g);
    int i, j, k;
    if (!region_point(idx, lo, hi, i, j, k)) return;
    if (d == 0)
        rhoh_face<0>(B, g, A, i, j, k);
    else if (d == 1)
        rhoh_face<1>(B, g, A, i, j, k);
    else
        rhoh_face<2>(B, g, A, i, j, k);
}

BaseP basep(const mx_basestate* s, int lev) {
    BaseP b;
    b.p = (s && s->p) ? s->p : nullptr;
    b.off = (s && s->p) ? (long long)lev * s->len : 0;
    return b;
}

DA face_da(const mx_face_multifab* f, int d, size_t b, bool use) {
    if (!use || !f || !f->dir[d].fab || (int)b >= f->dir[d].nfab) return null_da();
    return to_da(f->dir[d].fab[b]);
}

}  // namespace

namespace mxh {

// Maestro::BCSetup (MaestroSetup.cpp:276-559)
void bc_setup_host(mx_adv_ctx* c) {
    const mx_adv_params& p = c->p;
    c->bcs_s.assign(p.nscal, mx_bcrec{{0, 0, 0}, {0, 0, 0}});
    c->bcs_u.assign(3, mx_bcrec{{0, 0, 0}, {0, 0, 0}});
    for (int dir = 0; dir < p.dim; ++dir) {
        for (int side = 0; side < 2; ++side) {
            const int phys = side == 0 ? p.phys_bc_lo[dir] : p.phys_bc_hi[dir];
            int u_all, u_dir, sc, pi;
            switch (phys) {
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
                    sc = pi = MX_BC_REFLECT_EVEN;
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
                default:  // interior (periodic): internal Dirichlet
                    u_all = u_dir = sc = pi = MX_BC_INT_DIR;
                    break;
            }
            for (int comp = 0; comp < p.dim; ++comp) {
                int v = comp == dir ? u_dir : u_all;
                (side == 0 ? c->bcs_u[comp].lo : c->bcs_u[comp].hi)[dir] = v;
            }
            for (int comp = 0; comp < p.nscal; ++comp) {
                int v = comp == p.comp_pi ? pi : sc;
                (side == 0 ? c->bcs_s[comp].lo : c->bcs_s[comp].hi)[dir] = v;
            }
        }
    }
}

int make_edge_scal(mx_adv_ctx* c, int lev, const mx_multifab* state,
                   const mx_face_multifab* sedge, const mx_face_multifab* umac,
                   const mx_multifab* force, int is_vel, const mx_bcrec* bcs, int nbccomp,
                   int start_scomp, int start_bccomp, int num_comp, int is_conservative) {
    (void)nbccomp;
    if (fused_edge_scal_ok(c, is_conservative))
        return make_edge_scal_fused(c, lev, state, sedge, umac, force, is_vel, bcs, start_scomp,
                                    start_bccomp, num_comp, is_conservative);
    const int dim = c->p.dim;
    const auto& grids = c->grids[lev];
    const int ppm = c->p.ppm_type;
    const bool tf = c->p.ppm_trace_forces == 1 && ppm != 0;
    auto grown1 = [&](HBox b) { return grow(b, dim, 1); };

    size_t need = (size_t)num_comp * 2 * tmp_bytes(grids, dim, grown1);
    if (tf) need += (size_t)num_comp * 2 * tmp_bytes(grids, dim, grown1);
    if (dim == 3) need += (size_t)num_comp * 6 * tmp_bytes(grids, 1, grown1);
    int rc = ws_reserve(c, need);
    if (rc) return rc;

    std::vector<EsTmp> tmps(grids.size() * num_comp);
    for (int vc = 0; vc < num_comp; ++vc) {
        auto Ip = tmp_fabs(c, grids, dim, grown1);
        auto Im = tmp_fabs(c, grids, dim, grown1);
        std::vector<DA> Ipf, Imf, T[6];
        if (tf) {
            Ipf = tmp_fabs(c, grids, dim, grown1);
            Imf = tmp_fabs(c, grids, dim, grown1);
        }
        if (dim == 3)
            for (int p = 0; p < 6; ++p) T[p] = tmp_fabs(c, grids, 1, grown1);
        for (size_t b = 0; b < grids.size(); ++b) {
            EsTmp& t = tmps[b * num_comp + vc];
            t.Ip = Ip[b];
            t.Im = Im[b];
            t.Ipf = tf ? Ipf[b] : null_da();
            t.Imf = tf ? Imf[b] : null_da();
            for (int p = 0; p < 6; ++p) t.T[p] = dim == 3 ? T[p][b] : null_da();
            for (int d = 0; d < 3; ++d) {
                t.bclo[d] = bcs[start_bccomp + vc].lo[d];
                t.bchi[d] = bcs[start_bccomp + vc].hi[d];
            }
        }
    }
    std::vector<EsBox> boxes;
    long long maxpts = 0;
    for (size_t b = 0; b < grids.size(); ++b) {
        EsBox B;
        HBox g = hbox(grids[b]);
        for (int q = 0; q < 3; ++q) {
            B.lo[q] = g.lo[q];
            B.hi[q] = g.hi[q];
        }
        B.s = to_da(state->fab[b]);
        B.fo = to_da(force->fab[b]);
        for (int d = 0; d < 3; ++d) {
            B.um[d] = d < dim ? to_da(umac->dir[d].fab[b]) : null_da();
            B.se[d] = d < dim ? to_da(sedge->dir[d].fab[b]) : null_da();
        }
        maxpts = std::max(maxpts, npts(grown1(g)));
        boxes.push_back(B);
    }
    EsArgs A;
    std::memset(&A, 0, sizeof(A));
    A.start_scomp = start_scomp;
    A.num_comp = num_comp;
    A.is_vel = is_vel;
    A.is_conservative = is_conservative;

    Table tab(c);
    size_t ob = tab.add(boxes);
    size_t ot = tab.add(tmps);
    rc = tab.upload();
    if (rc) return rc;
    const EsBox* db = tab.dev<EsBox>(ob);
    const EsTmp* dt = tab.dev<EsTmp>(ot);
    KGeom kg = kgeom(c, lev);
    const int block = 256;
    const int nb = (int)boxes.size();

    dim3 gr = grid_for(maxpts, nb, num_comp, block);
    if (ppm == 0)
        MX_LAUNCH(c, "k_es_ppm", k_es_ppm<0>,  gr, dim3(block), db, dt, kg, A);
    else if (ppm == 1)
        MX_LAUNCH(c, "k_es_ppm", k_es_ppm<1>,  gr, dim3(block), db, dt, kg, A);
    else
        MX_LAUNCH(c, "k_es_ppm", k_es_ppm<2>,  gr, dim3(block), db, dt, kg, A);

    if (dim == 3) {
        MX_LAUNCH(c, "k_es_trans", k_es_trans,  grid_for(maxpts, nb, 6 * num_comp, block), dim3(block), db, dt, kg, A);

    }
    MX_LAUNCH(c, "k_es_edge", k_es_edge,  grid_for(maxpts, nb, dim * num_comp, block), dim3(block), db, dt, kg, A);

    return MX_OK;
}

static int flux_common(mx_adv_ctx* c, int lev, const mx_face_multifab* sflux,
                       const mx_multifab* etarhoflux, const mx_face_multifab* sedge,
                       const mx_face_multifab* umac, const mx_face_multifab* r0mac_old,
                       const mx_face_multifab* r0mac_new, const mx_face_multifab* rh0mac_old,
                       const mx_face_multifab* rh0mac_new, const mx_face_multifab* h0mac_old,
                       const mx_face_multifab* h0mac_new, std::vector<FluxBox>& boxes,
                       long long& maxpts) {
    const int dim = c->p.dim;
    const auto& grids = c->grids[lev];
    const bool sph = dim == 3 && c->p.spherical;
    maxpts = 0;
    for (size_t b = 0; b < grids.size(); ++b) {
        FluxBox B;
        HBox g = hbox(grids[b]);
        for (int q = 0; q < 3; ++q) {
            B.lo[q] = g.lo[q];
            B.hi[q] = g.hi[q];
        }
        for (int d = 0; d < 3; ++d) {
            const bool on = d < dim;
            B.sf[d] = face_da(sflux, d, b, on);
            B.se[d] = face_da(sedge, d, b, on);
            B.um[d] = face_da(umac, d, b, on);
            B.r0o[d] = face_da(r0mac_old, d, b, on && sph);
            B.r0n[d] = face_da(r0mac_new, d, b, on && sph);
            B.rh0o[d] = face_da(rh0mac_old, d, b, on && sph);
            B.rh0n[d] = face_da(rh0mac_new, d, b, on && sph);
            B.h0o[d] = face_da(h0mac_old, d, b, on && sph);
            B.h0n[d] = face_da(h0mac_new, d, b, on && sph);
        }
        B.eta = (etarhoflux && etarhoflux->fab) ? to_da(etarhoflux->fab[b]) : null_da();
        for (int d = 0; d < dim; ++d) maxpts = std::max(maxpts, npts(nodal(g, d)));
        boxes.push_back(B);
    }
    return MX_OK;
}

static FluxArgs flux_args(const mx_adv_ctx* c, int lev) {
    FluxArgs A;
    std::memset(&A, 0, sizeof(A));
    A.Rho = c->p.comp_rho;
    A.RhoH = c->p.comp_rhoh;
    A.FirstSpec = c->p.first_spec;
    A.NumSpec = c->p.num_spec;
    A.species_pred_type = c->p.species_pred_type;
    A.enthalpy_pred_type = c->p.enthalpy_pred_type;
    A.spherical = c->p.spherical;
    A.use_exact = c->p.use_exact_base_state;
    A.evolve = c->p.evolve_base_state;
    (void)lev;
    return A;
}

int make_rhox_flux(mx_adv_ctx* c, int lev, const mx_face_multifab* sflux,
                   const mx_multifab* etarhoflux, const mx_face_multifab* sedge,
                   const mx_face_multifab* umac, const mx_basestate* rho0_old,
                   const mx_basestate* rho0_edge_old, const mx_face_multifab* r0mac_old,
                   const mx_basestate* rho0_new, const mx_basestate* rho0_edge_new,
                   const mx_face_multifab* r0mac_new, const mx_basestate* rho0_predicted_edge,
                   const mx_basestate* w0, int start_comp, int num_comp, int zero_eta) {
    std::vector<FluxBox> boxes;
    long long maxpts;
    flux_common(c, lev, sflux, etarhoflux, sedge, umac, r0mac_old, r0mac_new, nullptr, nullptr,
                nullptr, nullptr, boxes, maxpts);
    FluxArgs A = flux_args(c, lev);
    A.rho0_old = basep(rho0_old, lev);
    A.rho0_edge_old = basep(rho0_edge_old, lev);
    A.rho0_new = basep(rho0_new, lev);
    A.rho0_edge_new = basep(rho0_edge_new, lev);
    A.rho0_pe = basep(rho0_predicted_edge, lev);
    A.w0 = basep(w0, lev);
    A.start_comp = start_comp;
    A.num_comp = num_comp;

    std::vector<FluxBox> zb;
    long long zmax = 0;
    if (zero_eta) {
        // etarhoflux[lev].setVal(0.): the whole FAB, ghosts included
        for (size_t b = 0; b < boxes.size(); ++b) {
            const mx_fab& f = etarhoflux->fab[b];
            FluxBox z = boxes[b];
            for (int q = 0; q < 3; ++q) {
                z.lo[q] = f.begin[q];
                z.hi[q] = f.end[q] - 1;
            }
            HBox r;
            for (int q = 0; q < 3; ++q) {
                r.lo[q] = z.lo[q];
                r.hi[q] = z.hi[q];
            }
            zmax = std::max(zmax, npts(r));
            zb.push_back(z);
        }
    }
    Table tab(c);
    size_t ob = tab.add(boxes);
    size_t oz = zero_eta ? tab.add(zb) : 0;
    int rc = tab.upload();
    if (rc) return rc;
    const int block = 256;
    const int nb = (int)boxes.size();
    if (zero_eta) {
        MX_LAUNCH(c, "k_zero_fab", k_zero_fab,  grid_for(zmax, nb, 1, block), dim3(block), tab.dev<FluxBox>(oz), nb);

    }
    const KGeom kg = kgeom(c, lev);
    MX_LAUNCH(c, "k_rhox_flux", k_rhox_flux, grid_for_g(kg, maxpts, nb, c->p.dim, block),
              dim3(block), tab.dev<FluxBox>(ob), kg, A);
    return flux_reg_ops(c, lev, sflux, true, start_comp, num_comp);
}

int make_rhoh_flux(mx_adv_ctx* c, int lev, const mx_face_multifab* sflux,
                   const mx_face_multifab* sedge, const mx_face_multifab* umac,
                   const mx_basestate* rho0_old, const mx_basestate* rho0_edge_old,
                   const mx_face_multifab* r0mac_old, const mx_basestate* rho0_new,
                   const mx_basestate* rho0_edge_new, const mx_face_multifab* r0mac_new,
                   const mx_basestate* rhoh0_old, const mx_basestate* rhoh0_edge_old,
                   const mx_face_multifab* rh0mac_old, const mx_basestate* rhoh0_new,
                   const mx_basestate* rhoh0_edge_new, const mx_face_multifab* rh0mac_new,
                   const mx_face_multifab* h0mac_old, const mx_face_multifab* h0mac_new) {
    std::vector<FluxBox> boxes;
    long long maxpts;
    flux_common(c, lev, sflux, nullptr, sedge, umac, r0mac_old, r0mac_new, rh0mac_old,
                rh0mac_new, h0mac_old, h0mac_new, boxes, maxpts);
    FluxArgs A = flux_args(c, lev);
    A.rho0_old = basep(rho0_old, lev);
    A.rho0_edge_old = basep(rho0_edge_old, lev);
    A.rho0_new = basep(rho0_new, lev);
    A.rho0_edge_new = basep(rho0_edge_new, lev);
    A.rhoh0_old = basep(rhoh0_old, lev);
    A.rhoh0_edge_old = basep(rhoh0_edge_old, lev);
    A.rhoh0_new = basep(rhoh0_new, lev);
    A.rhoh0_edge_new = basep(rhoh0_edge_new, lev);
    Table tab(c);
    size_t ob = tab.add(boxes);
    int rc = tab.upload();
    if (rc) return rc;
    const int block = 256;
    const KGeom kg = kgeom(c, lev);
    MX_LAUNCH(c, "k_rhoh_flux", k_rhoh_flux,
              grid_for_g(kg, maxpts, (int)boxes.size(), c->p.dim, block), dim3(block),
              tab.dev<FluxBox>(ob), kg, A);
    return flux_reg_ops(c, lev, sflux, false, 0, 0);
}

}  // namespace mxh

// ---------------------------------------------------------------------------
// C-ABI

namespace {

bool base_ok(const mx_basestate* s, int lev, int need_len) {
    return s && s->p && s->len >= need_len && s->nlev > lev;
}

}  // namespace

namespace mxh {

// argument checks of the pass on one level (mx_adv_advance_pass)
int pass_check(mx_adv_ctx* c, int lev, const mx_adv_pass_args* a) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    if (!a) return MX_ERR_ARG;
    const mx_adv_params& p = c->p;
    const bool sph = p.dim == 3 && p.spherical;
    if (sph) {
        if ((rc = check_face(c, lev, &a->w0mac, 1, 0))) return rc;
        if ((rc = check_face(c, lev, &a->r0mac, 1, 0))) return rc;
        if (p.use_exact_base_state) {
            if ((rc = check_face(c, lev, &a->rh0mac, 1, 0))) return rc;
        } else {
            if ((rc = check_face(c, lev, &a->h0mac, 1, 0))) return rc;
        }
    }
    const int dim = p.dim;
    const int g = p.ppm_type == 2 ? 4 : 3;
    const int fg = p.ppm_trace_forces == 1 ? g : 1;
    const int nr = p.domain_hi[lev][dim - 1] + 1;
    if ((rc = check_mf(c, lev, &a->utilde, dim, g, -1))) return rc;
    if ((rc = check_mf(c, lev, &a->ufull, dim, 1, -1))) return rc;
    if ((rc = check_mf(c, lev, &a->vel_force, dim, fg, -1))) return rc;
    if ((rc = check_mf(c, lev, &a->w0_cart, dim, sph ? 0 : 1, -1))) return rc;
    if ((rc = check_mf(c, lev, &a->scal, p.nscal, g, -1))) return rc;
    if ((rc = check_mf(c, lev, &a->scal_force, p.nscal, fg, -1))) return rc;
    if ((rc = check_face(c, lev, &a->utrans, 1, 1))) return rc;
    if ((rc = check_face(c, lev, &a->umac, 1, 1))) return rc;
    if ((rc = check_face(c, lev, &a->sedge, p.nscal, 0))) return rc;
    if ((rc = check_face(c, lev, &a->sflux, p.nscal, 0))) return rc;
    if ((rc = check_mf(c, lev, &a->etarhoflux, 1, 0, dim - 1))) return rc;
    if (!base_ok(&a->w0, lev, nr + 1) || !base_ok(&a->rho0, lev, nr) ||
        !base_ok(&a->rho0_edge, lev, nr + 1) || !base_ok(&a->rhoh0, lev, nr) ||
        !base_ok(&a->rhoh0_edge, lev, nr + 1) || !base_ok(&a->rho0_predicted_edge, lev, nr + 1))
        return MX_ERR_ARG;
    if (dim == 2 && p.ppm_trace_forces == 1) return MX_ERR_UNSUPPORTED;
    return MX_OK;
}

// The pass over levels lev0 .. lev0 + nlev - 1 (a[i] holds level lev0 + i).
// Levels are independent, so the velocity stages run level by level, the
// scalar edge states of every level go to one fused launch per component
// group (MakeEdgeScal's species, Rho, RhoH calls) and the fluxes follow.
int advance_pass_levels(mx_adv_ctx* c, int lev0, int nlev, const mx_adv_pass_args* a) {
    if (!c || !a || nlev < 1 || lev0 < 0 || lev0 + nlev - 1 > c->p.finest_level)
        return MX_ERR_ARG;
    int rc;
    for (int i = 0; i < nlev; ++i)
        if ((rc = pass_check(c, lev0 + i, &a[i]))) return rc;
    const mx_adv_params& p = c->p;
    const bool sph = p.dim == 3 && p.spherical;
    const mx_bcrec* bs = c->bcs_s.data();
    const int nbs = (int)c->bcs_s.size();
    // a hierarchy (levels 0..nlev-1, nlev > 1) synchronises its face data
    // between levels; a single level fills its ghost faces on its own
    const bool multi = nlev > 1;
    if (multi && lev0 != 0) return MX_ERR_ARG;
    std::vector<VpPlan> vps(nlev);
    // the fused VelPred kernel keeps its temporaries on chip: MakeUtrans
    // reconstructs its own normal states and VelPred follows in one launch
    const bool vfused = velpred_fused_ok(c);
    std::vector<VpLevel> vlevs;
    for (int i = 0; i < nlev && vfused; ++i) {
        const mx_adv_pass_args* A = &a[i];
        vlevs.push_back({lev0 + i, &A->utilde, &A->ufull, &A->utrans, &A->umac, &A->w0mac,
                         &A->vel_force, &A->w0_cart});
    }
    if (multi && !vfused) {
        // every level's VelPred temporaries stay live until the levels'
        // utrans are synchronised
        size_t need = 0;
        for (int i = 0; i < nlev; ++i) need += velpred_ws_bytes(c, lev0 + i);
        if ((rc = ws_reserve(c, need))) return rc;
    }
    for (int i = 0; i < nlev && vfused; ++i) {
        const int lev = lev0 + i;
        const mx_adv_pass_args* A = &a[i];
        if ((rc = make_utrans(c, lev, &A->utilde, &A->ufull, &A->utrans, &A->w0mac, &A->w0_cart,
                              !multi)))
            return rc;
        if ((rc = addw0(c, lev, &A->utrans, &A->w0mac, &A->w0, 1.0, !multi))) return rc;
        if (multi) continue;
        std::vector<VpLevel> one(1, vlevs[i]);
        if ((rc = velpred_fused(c, one))) return rc;
        if ((rc = addw0(c, lev, &A->umac, &A->w0mac, &A->w0, 1.0))) return rc;
    }
    for (int i = 0; i < nlev && !vfused; ++i) {
        const int lev = lev0 + i;
        const mx_adv_pass_args* A = &a[i];
        // VelPred's reconstruction runs first and MakeUtrans reads its traced
        // normal-direction states (the numbers MakeUtrans would compute)
        VpPlan& vp = vps[i];
        if ((rc = velpred_setup(c, lev, &A->utilde, &A->ufull, &A->utrans, &A->umac, &A->w0mac,
                                &A->vel_force, &A->w0_cart, vp, !multi)))
            return rc;
        if ((rc = utrans_from_plan(c, vp, !multi))) return rc;
        if ((rc = addw0(c, lev, &A->utrans, &A->w0mac, &A->w0, 1.0, !multi))) return rc;
        if (multi) continue;
        if ((rc = velpred_finish(c, vp))) return rc;
        if ((rc = addw0(c, lev, &A->umac, &A->w0mac, &A->w0, 1.0))) return rc;
    }
    // AverageDownFaces over the hierarchy, then FillPatchUedge
    auto sync_levels = [&](mx_face_multifab mx_adv_pass_args::*field, bool fill) -> int {
        std::vector<mx_face_multifab> f(nlev);
        for (int i = 0; i < nlev; ++i) f[i] = a[i].*field;
        for (int lev = nlev - 2; lev >= 0; --lev) {
            const int r = mx_adv_average_down_faces(c, lev, &f[lev + 1], &f[lev]);
            if (r) return r;
        }
        return fill ? fill_patch_uedge(c, nlev, f.data()) : MX_OK;
    };
    if (multi) {
        // The reference synchronises utrans after MakeUtrans and again after
        // Addw0 (MaestroMakeUtrans.cpp:641-647, MaestroFill3dData.cpp:449-455),
        // and umac after VelPred and again after Addw0 (MaestroVelPred.cpp:291).
        // Both syncs read valid faces only and the second rewrites every face
        // the first one wrote, so one sync after the w0 addition gives the
        // same values.
        if ((rc = sync_levels(&mx_adv_pass_args::utrans, true))) return rc;
        if (vfused && (rc = velpred_fused(c, vlevs))) return rc;
        for (int i = 0; i < nlev; ++i) {
            const mx_adv_pass_args* A = &a[i];
            if (!vfused && (rc = velpred_finish(c, vps[i]))) return rc;
            if ((rc = addw0(c, lev0 + i, &A->umac, &A->w0mac, &A->w0, 1.0, false))) return rc;
        }
        if ((rc = sync_levels(&mx_adv_pass_args::umac, true))) return rc;
    }
    const int groups[3][2] = {{p.first_spec, p.num_spec}, {p.comp_rho, 1}, {p.comp_rhoh, 1}};
    // MakeRhoHFlux's planar branches that read only the RhoH edge state run
    // as the fused edge kernel's epilogue (the flux from the edge state in
    // registers); the register updates of reflux_type 2 follow as usual
    const int ept = p.enthalpy_pred_type;
    const bool rhoh_epilogue = fused_edge_scal_ok(c, 0) && c->rhoh_flux_epilogue && !sph &&
                               (ept == 0 || ept == 1 || ept == 3);
    if (fused_edge_scal_ok(c, 0)) {
        std::vector<EdgeLevel> levs;
        for (int i = 0; i < nlev; ++i) {
            EdgeLevel L{lev0 + i, &a[i].scal, &a[i].sedge, &a[i].umac, &a[i].scal_force};
            if (rhoh_epilogue) {
                const long long o = (long long)(lev0 + i) * a[i].rhoh0.len;
                const long long oe = (long long)(lev0 + i) * a[i].rhoh0_edge.len;
                L.flux_comp = p.comp_rhoh;
                L.flux_kind = ept == 0 ? 2 : 1;
                L.sflux = &a[i].sflux;
                L.bc_a = L.bc_b = a[i].rhoh0.p + o;       // old and new: the pass's one state
                L.be_a = L.be_b = a[i].rhoh0_edge.p + oe;
            }
            levs.push_back(L);
        }
        // MakeEdgeScal's three calls (species, Rho, RhoH) read the same
        // inputs and write disjoint components: one launch covers them all
        std::vector<int> sc;
        for (const auto& gr : groups)
            for (int n = 0; n < gr[1]; ++n) sc.push_back(gr[0] + n);
        if ((rc = make_edge_scal_fused_list(c, levs, 0, bs, sc, sc, 0))) return rc;
    } else {
        for (int i = 0; i < nlev; ++i)
            for (const auto& gr : groups)
                if ((rc = make_edge_scal(c, lev0 + i, &a[i].scal, &a[i].sedge, &a[i].umac,
                                         &a[i].scal_force, 0, bs, nbs, gr[0], gr[0], gr[1], 0)))
                    return rc;
    }
    for (int i = 0; i < nlev; ++i) {
        const int lev = lev0 + i;
        const mx_adv_pass_args* A = &a[i];
        const mx_face_multifab* r0m = sph ? &A->r0mac : nullptr;
        const mx_face_multifab* rh0m = sph ? &A->rh0mac : nullptr;
        const mx_face_multifab* h0m = sph ? &A->h0mac : nullptr;
        if ((rc = make_rhox_flux(c, lev, &A->sflux, &A->etarhoflux, &A->sedge, &A->umac,
                                 &A->rho0, &A->rho0_edge, r0m, &A->rho0, &A->rho0_edge, r0m,
                                 &A->rho0_predicted_edge, &A->w0, p.first_spec, p.num_spec, 1)))
            return rc;
        if (rhoh_epilogue) {
            if ((rc = flux_reg_ops(c, lev, &A->sflux, false, 0, 0))) return rc;
        } else if ((rc = make_rhoh_flux(c, lev, &A->sflux, &A->sedge, &A->umac, &A->rho0,
                                        &A->rho0_edge, r0m, &A->rho0, &A->rho0_edge, r0m,
                                        &A->rhoh0, &A->rhoh0_edge, rh0m, &A->rhoh0,
                                        &A->rhoh0_edge, rh0m, h0m, h0m))) {
            return rc;
        }
    }
    // reflux_type 1: AverageDownFaces(sflux) after MakeRhoXFlux and after
    // MakeRhoHFlux (MaestroMakeFlux.cpp:459-462, 1061-1064).  MakeRhoHFlux
    // writes only the RhoH component and reads no flux, so one restriction
    // of every component at the end gives the same values.  reflux_type 2
    // has updated the flux registers inside the flux stages instead.
    if (multi && p.reflux_type == 1 && (rc = sync_levels(&mx_adv_pass_args::sflux, false)))
        return rc;
    return MX_OK;
}

}  // namespace mxh

extern "C" {

int mx_adv_bc_setup(const mx_adv_ctx* c, mx_bcrec* bcs_s, mx_bcrec* bcs_u) {
    if (!c) return MX_ERR_ARG;
    if (bcs_s) std::memcpy(bcs_s, c->bcs_s.data(), sizeof(mx_bcrec) * c->bcs_s.size());
    if (bcs_u) std::memcpy(bcs_u, c->bcs_u.data(), sizeof(mx_bcrec) * c->p.dim);
    return MX_OK;
}

int mx_adv_make_edge_scal(mx_adv_ctx* c, int lev, const mx_multifab* state,
                          mx_face_multifab* sedge, const mx_face_multifab* umac,
                          const mx_multifab* force, int is_vel, const mx_bcrec* bcs, int nbccomp,
                          int start_scomp, int start_bccomp, int num_comp, int is_conservative) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    if (!bcs || num_comp <= 0 || num_comp > MX_MAX_SCAL || start_scomp < 0 || start_bccomp < 0 ||
        start_bccomp + num_comp > nbccomp)
        return MX_ERR_ARG;
    const int dim = c->p.dim;
    const int g = c->p.ppm_type == 2 ? 4 : 3;
    const int nc = start_scomp + num_comp;
    if ((rc = check_mf(c, lev, state, nc, g, -1))) return rc;
    if ((rc = check_face(c, lev, sedge, nc, 0))) return rc;
    if ((rc = check_face(c, lev, umac, 1, 1))) return rc;
    const int fg = c->p.ppm_trace_forces == 1 ? g : 1;
    if ((rc = check_mf(c, lev, force, nc, fg, -1))) return rc;
    (void)dim;
    return make_edge_scal(c, lev, state, sedge, umac, force, is_vel, bcs, nbccomp, start_scomp,
                          start_bccomp, num_comp, is_conservative);
}

int mx_adv_make_rhox_flux(mx_adv_ctx* c, int lev, mx_face_multifab* sflux,
                          mx_multifab* etarhoflux, const mx_face_multifab* sedge,
                          const mx_face_multifab* umac, const mx_basestate* rho0_old,
                          const mx_basestate* rho0_edge_old, const mx_face_multifab* r0mac_old,
                          const mx_basestate* rho0_new, const mx_basestate* rho0_edge_new,
                          const mx_face_multifab* r0mac_new,
                          const mx_basestate* rho0_predicted_edge, const mx_basestate* w0,
                          int start_comp, int num_comp) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    const int dim = c->p.dim;
    const int nscal = c->p.nscal;
    if (start_comp < 0 || num_comp < 0 || start_comp + num_comp > nscal) return MX_ERR_ARG;
    if ((rc = check_face(c, lev, sflux, nscal, 0))) return rc;
    if ((rc = check_face(c, lev, sedge, nscal, 0))) return rc;
    if ((rc = check_face(c, lev, umac, 1, 0))) return rc;
    const int nr = c->p.domain_hi[lev][dim - 1] + 1;
    if (dim == 3 && c->p.spherical) {
        if ((rc = check_face(c, lev, r0mac_old, 1, 0))) return rc;
        if ((rc = check_face(c, lev, r0mac_new, 1, 0))) return rc;
    } else {
        if (!base_ok(rho0_old, lev, nr) || !base_ok(rho0_new, lev, nr) ||
            !base_ok(rho0_edge_old, lev, nr + 1) || !base_ok(rho0_edge_new, lev, nr + 1))
            return MX_ERR_ARG;
        if (c->p.evolve_base_state && !c->p.use_exact_base_state) {
            if (!base_ok(rho0_predicted_edge, lev, nr + 1) || !base_ok(w0, lev, nr + 1))
                return MX_ERR_ARG;
            if ((rc = check_mf(c, lev, etarhoflux, 1, 0, dim - 1))) return rc;
        }
    }
    return make_rhox_flux(c, lev, sflux, etarhoflux, sedge, umac, rho0_old, rho0_edge_old,
                          r0mac_old, rho0_new, rho0_edge_new, r0mac_new, rho0_predicted_edge, w0,
                          start_comp, num_comp, 0);
}

int mx_adv_make_rhoh_flux(mx_adv_ctx* c, int lev, mx_face_multifab* sflux,
                          const mx_face_multifab* sedge, const mx_face_multifab* umac,
                          const mx_face_multifab* w0mac, const mx_basestate* rho0_old,
                          const mx_basestate* rho0_edge_old, const mx_face_multifab* r0mac_old,
                          const mx_basestate* rho0_new, const mx_basestate* rho0_edge_new,
                          const mx_face_multifab* r0mac_new, const mx_basestate* rhoh0_old,
                          const mx_basestate* rhoh0_edge_old, const mx_face_multifab* rh0mac_old,
                          const mx_basestate* rhoh0_new, const mx_basestate* rhoh0_edge_new,
                          const mx_face_multifab* rh0mac_new, const mx_face_multifab* h0mac_old,
                          const mx_face_multifab* h0mac_new) {
    (void)w0mac;
    int rc = check_level(c, lev);
    if (rc) return rc;
    const int dim = c->p.dim;
    const int nscal = c->p.nscal;
    if ((rc = check_face(c, lev, sflux, nscal, 0))) return rc;
    if ((rc = check_face(c, lev, sedge, nscal, 0))) return rc;
    if ((rc = check_face(c, lev, umac, 1, 0))) return rc;
    const int nr = c->p.domain_hi[lev][dim - 1] + 1;
    if (dim == 3 && c->p.spherical) {
        if (c->p.use_exact_base_state) {
            if ((rc = check_face(c, lev, rh0mac_old, 1, 0))) return rc;
            if ((rc = check_face(c, lev, rh0mac_new, 1, 0))) return rc;
        } else {
            if ((rc = check_face(c, lev, r0mac_old, 1, 0))) return rc;
            if ((rc = check_face(c, lev, r0mac_new, 1, 0))) return rc;
            if ((rc = check_face(c, lev, h0mac_old, 1, 0))) return rc;
            if ((rc = check_face(c, lev, h0mac_new, 1, 0))) return rc;
        }
    } else {
        if (!base_ok(rho0_old, lev, nr) || !base_ok(rho0_new, lev, nr) ||
            !base_ok(rho0_edge_old, lev, nr + 1) || !base_ok(rho0_edge_new, lev, nr + 1) ||
            !base_ok(rhoh0_old, lev, nr) || !base_ok(rhoh0_new, lev, nr) ||
            !base_ok(rhoh0_edge_old, lev, nr + 1) || !base_ok(rhoh0_edge_new, lev, nr + 1))
            return MX_ERR_ARG;
    }
    return make_rhoh_flux(c, lev, sflux, sedge, umac, rho0_old, rho0_edge_old, r0mac_old,
                          rho0_new, rho0_edge_new, r0mac_new, rhoh0_old, rhoh0_edge_old,
                          rh0mac_old, rhoh0_new, rhoh0_edge_new, rh0mac_new, h0mac_old,
                          h0mac_new);
}

int mx_adv_advance_pass(mx_adv_ctx* c, int lev, const mx_adv_pass_args* a) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    return advance_pass_levels(c, lev, 1, a);
}

}  // extern "C"
