// mx_velocity.hip -- MakeUtrans and VelPred on the GPU.
//
// MakeUtrans (MaestroMakeUtrans.cpp:6-648): one kernel over the faces of
// every box and direction.  Each face thread reconstructs the normal velocity
// component in the normal direction only, on its two adjacent cells (the
// reference runs PPM in all three directions and keeps one), applies the
// physical-boundary overrides and solves the Riemann problem, with the
// cell-centred w0 in the vertical direction (planar) or w0mac (spherical).
//
// VelPred (MaestroVelPred.cpp:5-1915): three kernels over all boxes --
//   k_vp_ppm     PPM (or slopes) of every velocity component in every
//                direction on the box grown by one (the Ipu..Imw temporaries);
//   k_vp_trans   the six transverse states uimhyz..wimhyx (3-D), with the
//                interface states ulx..urz and the upwind states uimhx..uimhz
//                evaluated on the fly from the PPM temporaries;
//   k_vp_vel     the MAC velocities umac/vmac/wmac on the valid faces.
#include <algorithm>

#include "mx_internal.H"

using namespace mxd;
using namespace mxh;

namespace {

constexpr int e(int d, int q) { return d == q ? 1 : 0; }

// index of the transverse state on d-faces corrected by t (d != t)
__host__ __device__ constexpr int pair_index(int d, int t) { return d * 2 + (t > d ? t - 1 : t); }

// ---------------------------------------------------------------------------
// MakeUtrans

struct UtBox {
    int lo[3], hi[3];
    DA ut, uf, w0c;
    DA tr[3];
    DA w0m[3];
};

struct UtArgs {
    int bclo[3][3], bchi[3][3];  // bcs_u[comp].lo/hi[dir]
};

template <int D, class BOX>
__device__ __forceinline__ void utrans_riemann(const BOX& B, const KGeom& g, int i, int j, int k,
                                               double ul, double ur);

template <int PPM, int D>
__device__ __forceinline__ void utrans_face(const UtBox& B, const KGeom& g, const UtArgs& A,
                                            int i, int j, int k) {
    const int im = i - e(D, 0), jm = j - e(D, 1), km = k - e(D, 2);
    const int pos = D == 0 ? i : (D == 1 ? j : k);
    const int posm = pos - 1;
    const double dt2 = 0.5 * g.dt;
    const double hd = g.dx[D];
    const int bclo = A.bclo[D][D], bchi = A.bchi[D][D];
    const long long st = B.ut.stride(D);
    double ul, ur;
    if (PPM == 0) {
        double v[9];
        load_stencil<1>(B.ut.p + B.ut.off(im, jm, km, D), st, v);
        double slL = slope_point<D>(v, g.slope_order, posm, g.domlo[D], g.domhi[D], bclo, bchi);
        load_stencil<1>(B.ut.p + B.ut.off(i, j, k, D), st, v);
        double slR = slope_point<D>(v, g.slope_order, pos, g.domlo[D], g.domhi[D], bclo, bchi);
        ul = B.ut(im, jm, km, D) + (0.5 - (dt2 / hd) * dmax(0.0, B.uf(im, jm, km, D))) * slL;
        ur = B.ut(i, j, k, D) - (0.5 + (dt2 / hd) * dmin(0.0, B.uf(i, j, k, D))) * slR;
    } else {
        double v[9], sm, sp, Ip, Im;
        load_stencil<PPM>(B.ut.p + B.ut.off(im, jm, km, D), st, v);
        ppm_edges<PPM, D>(v, posm, g.domlo[D], g.domhi[D], bclo, bchi, sm, sp);
        double uc = B.uf(im, jm, km, D);
        ppm_trace(v[4], sm, sp, uc, uc, g.dt, hd, g.rel_eps, Ip, Im);
        ul = Ip;
        load_stencil<PPM>(B.ut.p + B.ut.off(i, j, k, D), st, v);
        ppm_edges<PPM, D>(v, pos, g.domlo[D], g.domhi[D], bclo, bchi, sm, sp);
        uc = B.uf(i, j, k, D);
        ppm_trace(v[4], sm, sp, uc, uc, g.dt, hd, g.rel_eps, Ip, Im);
        ur = Im;
    }

    utrans_riemann<D>(B, g, i, j, k, ul, ur);
}

// the PhysBC overrides and the Riemann solve of MakeUtrans on D-face
// (i, j, k) from its left and right states (MaestroMakeUtrans.cpp:109-150,
// 253-266); BOX: any box descriptor with ut, w0c, w0m and tr views
template <int D, class BOX>
__device__ __forceinline__ void utrans_riemann(const BOX& B, const KGeom& g, int i, int j, int k,
                                               double ul, double ur) {
    const int im = i - e(D, 0), jm = j - e(D, 1), km = k - e(D, 2);
    const int pos = D == 0 ? i : (D == 1 ? j : k);
    if (pos == g.domlo[D]) {
        switch (g.phys_lo[D]) {
            case MX_PHYS_INFLOW:
                ul = B.ut(im, jm, km, D);
                ur = B.ut(im, jm, km, D);
                break;
            case MX_PHYS_SLIPWALL:
            case MX_PHYS_NOSLIPWALL:
            case MX_PHYS_SYMMETRY:
                if (g.dim == 2 && D == 1) {
                    ur = 0.0;  // the 2-D y-lo branch zeroes vry twice and keeps vly
                } else {
                    ul = 0.0;
                    ur = 0.0;
                }
                break;
            case MX_PHYS_OUTFLOW:
                ul = dmin(ur, 0.0);
                ur = ul;
                break;
            default:
                break;
        }
    } else if (pos == g.domhi[D] + 1) {
        switch (g.phys_hi[D]) {
            case MX_PHYS_INFLOW:
                if (g.dim == 2) {
                    ul = B.ut(i, j, k, D);
                    ur = B.ut(i, j, k, D);
                } else {
                    ul = B.ut(i + e(D, 0), j + e(D, 1), k + e(D, 2), D);
                    ur = B.ut(i + e(D, 0), j + e(D, 1), k + e(D, 2), D);
                }
                break;
            case MX_PHYS_SLIPWALL:
            case MX_PHYS_NOSLIPWALL:
            case MX_PHYS_SYMMETRY:
                ul = 0.0;
                ur = 0.0;
                break;
            case MX_PHYS_OUTFLOW:
                ul = dmax(ul, 0.0);
                ur = ul;
                break;
            default:
                break;
        }
    }

    double res;
    const bool sph = g.dim == 3 && g.spherical;
    if (sph || D == g.dim - 1) {
        const double w = sph ? B.w0m[D](i, j, k) : B.w0c(i, j, k, g.dim - 1);
        bool test = (ul + w <= 0.0 && ur + w >= 0.0) || (dabs(ul + ur + 2.0 * w) < g.rel_eps);
        res = 0.5 * (ul + ur) + w > 0.0 ? ul : ur;
        res = test ? 0.0 : res;
    } else {
        bool test = (ul <= 0.0 && ur >= 0.0) || (dabs(ul + ur) < g.rel_eps);
        res = 0.5 * (ul + ur) > 0.0 ? ul : ur;
        res = test ? 0.0 : res;
    }
    B.tr[D](i, j, k) = res;
}

template <int PPM>
__global__ void __launch_bounds__(256) k_utrans(const UtBox* __restrict__ boxes, KGeom g, UtArgs A) {
    const UtBox& B = boxes[blockIdx.y];
    const int d = blockIdx.z;
    int lo[3] = {B.lo[0], B.lo[1], B.lo[2]};
    int hi[3] = {B.hi[0], B.hi[1], B.hi[2]};
    hi[d] += 1;
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!region_point(idx, lo, hi, i, j, k)) return;
    if (d == 0)
        utrans_face<PPM, 0>(B, g, A, i, j, k);
    else if (d == 1)
        utrans_face<PPM, 1>(B, g, A, i, j, k);
    else
        utrans_face<PPM, 2>(B, g, A, i, j, k);
}

// MakeUtrans with each cell traced once (3-D, ppm_type 1 and 2).  The state
// on a D-face depends only on velocity component D along direction D: the
// face takes Ip of its low cell and Im of its high cell, the same
// ppm_edges / ppm_trace evaluations that utrans_face makes per face, so the
// results are bit-identical.
//   blockIdx.z = 0, x-faces: a wavefront's 64 lanes take consecutive cells
//     of the box's rows, each row with its cells lo - 1 .. hi + 1 (nx + 2
//     cells, flattened; lanes consecutive in x).  Each lane traces its cell,
//     and the face at the low side of a lane's cell takes the previous
//     lane's Ip by a wavefront shuffle (no LDS, no barrier).  Wavefront w
//     starts at cell 63 w, so its last cell is the next one's first, whose
//     face the earlier wavefront forms.
//   blockIdx.z = 1, 2: each thread takes one x column (lanes consecutive in
//     x, so loads and stores coalesce) and one of nch chunks of the faces
//     lo .. hi + 1 along y or z, marching a register window along its column
//     and tracing each cell once (one extra cell per chunk).  The march also
//     carries the slopes (ppm_type 1) or face values (ppm_type 2) from cell
//     to cell, each evaluated once as in the fused VelPred's z direction, and
//     takes ppm_edges in full where the cell has a boundary stencil.
constexpr int UT_NT = 256;

template <int PPM>
__device__ __forceinline__ void utrans_rows(const UtBox& B, const KGeom& g, const UtArgs& A) {
    const unsigned nx = (unsigned)(B.hi[0] - B.lo[0] + 1);
    const unsigned ny = (unsigned)(B.hi[1] - B.lo[1] + 1);
    const unsigned nz = (unsigned)(B.hi[2] - B.lo[2] + 1);
    const unsigned rw = nx + 2;
    const unsigned lane = threadIdx.x & 63u;
    const unsigned wave = (blockIdx.x * (unsigned)UT_NT + threadIdx.x) >> 6;
    const unsigned t = wave * 63u + lane;
    const bool on = t < rw * ny * nz;
    unsigned c = 0;
    int i = 0, j = 0, k = 0;
    double Ip = 0.0, Im = 0.0;
    if (on) {
        const unsigned row = t / rw;
        c = t - row * rw;
        i = B.lo[0] - 1 + (int)c;
        j = B.lo[1] + (int)(row % ny);
        k = B.lo[2] + (int)(row / ny);
        double v[9], sm, sp;
        load_stencil<PPM>(B.ut.p + B.ut.off(i, j, k, 0), 1, v);
        ppm_edges<PPM, 0>(v, i, g.domlo[0], g.domhi[0], A.bclo[0][0], A.bchi[0][0], sm, sp);
        const double uc = B.uf(i, j, k, 0);
        ppm_trace(v[4], sm, sp, uc, uc, g.dt, g.dx[0], g.rel_eps, Ip, Im);
    }
    // every lane takes part in the shuffle; lane 0's result is unused
    const double Ipl = __shfl_up(Ip, 1u, 64);
    if (on && lane > 0 && c > 0) utrans_riemann<0>(B, g, i, j, k, Ipl, Im);
}

template <int PPM, int D>
__device__ __forceinline__ void utrans_march(const UtBox& B, const KGeom& g, const UtArgs& A,
                                             int nch) {
    constexpr int G = PPM == 2 ? 3 : 2;
    constexpr int T = D == 1 ? 2 : 1;  // the transverse direction other than x
    const unsigned nx = (unsigned)(B.hi[0] - B.lo[0] + 1);
    const unsigned nt = (unsigned)(B.hi[T] - B.lo[T] + 1);
    const int L = (B.hi[D] - B.lo[D] + 2 + nch - 1) / nch;
    const unsigned t = blockIdx.x * (unsigned)UT_NT + threadIdx.x;
    if (t >= nx * nt * (unsigned)nch) return;
    const unsigned r = t / nx;
    const int i = B.lo[0] + (int)(t - r * nx);
    const int tt = B.lo[T] + (int)(r % nt);
    const int f0 = B.lo[D] + (int)(r / nt) * L;
    const int f1 = min(f0 + L - 1, B.hi[D] + 1);
    if (f0 > f1) return;
    const long long st = B.ut.stride(D), fst = B.uf.stride(D);
    const double* up = B.ut.p + B.ut.off(i, D == 1 ? f0 - 1 : tt, D == 2 ? f0 - 1 : tt, D);
    const double* fp = B.uf.p + B.uf.off(i, D == 1 ? f0 - 1 : tt, D == 2 ? f0 - 1 : tt, D);
    const int dlo = g.domlo[D], dhi = g.domhi[D];
    const int bclo = A.bclo[D][D], bchi = A.bchi[D][D];
    double v[9];
    load_stencil<PPM>(up, st, v);
    double uc = *fp;
    double Ipl = 0.0;
    double zf0 = 0.0, zf1 = 0.0, zf2 = 0.0, zsl = 0.0;
    for (int cpos = f0 - 1;; ++cpos) {
        // the next cell's window entry and velocity, loaded ahead
        double vn = 0.0, un = 0.0;
        if (cpos < f1) {
            vn = up[(G + 1) * st];
            un = fp[fst];
        }
        double sm, sp, Ip, Im;
        if (PPM == 1) {
            if (cpos == f0 - 1) {
                const double d0 = vl_slope(v[2], v[3], v[4]);
                zsl = vl_slope(v[3], v[4], v[5]);
                zf0 = face1(v[3], v[4], d0, zsl);
            }
            const double d1 = vl_slope(v[4], v[5], v[6]);
            zf1 = face1(v[4], v[5], zsl, d1);
            zsl = d1;
            sm = zf0;
            sp = zf1;
            cw_limit(v[4], sm, sp);
        } else {
            if (cpos == f0 - 1) {
                zf0 = cs_face(v[1], v[2], v[3], v[4]);
                zf1 = cs_face(v[2], v[3], v[4], v[5]);
                zf2 = cs_face(v[3], v[4], v[5], v[6]);
            }
            const double f3 = cs_face(v[4], v[5], v[6], v[7]);
            cs_cell(v, zf0, zf1, zf2, f3, 1.e-10, sm, sp);
            zf0 = zf1;
            zf1 = zf2;
            zf2 = f3;
        }
        if (ppm_special<PPM>(cpos, dlo, dhi, bclo, bchi))
            ppm_edges<PPM, D>(v, cpos, dlo, dhi, bclo, bchi, sm, sp);
        if (PPM == 1) zf0 = zf1;
        ppm_trace(v[4], sm, sp, uc, uc, g.dt, g.dx[D], g.rel_eps, Ip, Im);
        if (cpos >= f0) utrans_riemann<D>(B, g, i, D == 1 ? cpos : tt, D == 2 ? cpos : tt, Ipl, Im);
        if (cpos == f1) break;
        Ipl = Ip;
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = v[q + 1];
        v[4 + G] = vn;
        uc = un;
        up += st;
        fp += fst;
    }
}

template <int PPM>
__global__ void __launch_bounds__(UT_NT) k_utrans_line(const UtBox* __restrict__ boxes, KGeom g,
                                                      UtArgs A, int nch) {
    const UtBox& B = boxes[blockIdx.y];
    if (blockIdx.z == 0)
        utrans_rows<PPM>(B, g, A);
    else if (blockIdx.z == 1)
        utrans_march<PPM, 1>(B, g, A, nch);
    else
        utrans_march<PPM, 2>(B, g, A, nch);
}

// ---------------------------------------------------------------------------
// VelPred

struct VpBox {
    int lo[3], hi[3];
    DA ut, uf, fo, w0c;
    DA tr[3], um[3], w0m[3];
    DA Ip[3], Im[3];    // per velocity component: [dir] comps; ppm_type 0: slopes along dir
    DA Ipf[3], Imf[3];  // traced forces
    DA T[6];            // transverse states, pair_index(d, t)
};

struct VpArgs {
    int bclo[3][3], bchi[3][3];  // bcs_u[comp].lo/hi[dir]
};

// the PPM traced states (or ppm_type 0: the slope) of velocity component n
// along D at cell (i,j,k); with traced forces also theirs (f0, f1)
template <int PPM, int D>
__device__ __forceinline__ void vp_ppm_dir(const VpBox& B, const KGeom& g, const VpArgs& A, int n,
                                           int i, int j, int k, double& r0, double& r1,
                                           double& f0, double& f1) {
    const int pos = D == 0 ? i : (D == 1 ? j : k);
    double v[9];
    if (PPM == 0) {
        load_stencil<1>(B.ut.p + B.ut.off(i, j, k, n), B.ut.stride(D), v);
        r0 = slope_point<D>(v, g.slope_order, pos, g.domlo[D], g.domhi[D], A.bclo[n][D],
                            A.bchi[n][D]);
        return;
    }
    double sm, sp;
    load_stencil<PPM>(B.ut.p + B.ut.off(i, j, k, n), B.ut.stride(D), v);
    ppm_edges<PPM, D>(v, pos, g.domlo[D], g.domhi[D], A.bclo[n][D], A.bchi[n][D], sm, sp);
    const double uc = B.uf(i, j, k, D);
    ppm_trace(v[4], sm, sp, uc, uc, g.dt, g.dx[D], g.rel_eps, r0, r1);
    if (g.ppm_trace_forces == 1) {
        load_stencil<PPM>(B.fo.p + B.fo.off(i, j, k, n), B.fo.stride(D), v);
        ppm_edges<PPM, D>(v, pos, g.domlo[D], g.domhi[D], A.bclo[n][D], A.bchi[n][D], sm, sp);
        ppm_trace(v[4], sm, sp, uc, uc, g.dt, g.dx[D], g.rel_eps, f0, f1);
    }
}

// every velocity component of a cell in one thread (ppm_type 1/2, no traced
// forces, 3-D): the tracing velocity and the index work are shared, and all
// loads come before the 18 stores
template <int PPM>
__global__ void __launch_bounds__(256) k_vp_ppm_all(const VpBox* __restrict__ boxes, KGeom g,
                                                    VpArgs A) {
    const VpBox& B = boxes[blockIdx.y];
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = B.lo[d] - 1;
        hi[d] = B.hi[d] + 1;
    }
    long long idx = grid_point(g);
    int i, j, k;
    if (!region_point(idx, lo, hi, i, j, k)) return;
    double r0[3][3], r1[3][3], f0, f1;
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        vp_ppm_dir<PPM, 0>(B, g, A, n, i, j, k, r0[n][0], r1[n][0], f0, f1);
        vp_ppm_dir<PPM, 1>(B, g, A, n, i, j, k, r0[n][1], r1[n][1], f0, f1);
        vp_ppm_dir<PPM, 2>(B, g, A, n, i, j, k, r0[n][2], r1[n][2], f0, f1);
    }
#pragma unroll
    for (int n = 0; n < 3; ++n)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            B.Ip[n](i, j, k, d) = r0[n][d];
            B.Im[n](i, j, k, d) = r1[n][d];
        }
}

template <int PPM>
__global__ void __launch_bounds__(256) k_vp_ppm(const VpBox* __restrict__ boxes, KGeom g, VpArgs A) {
    const VpBox& B = boxes[blockIdx.y];
    const int n = blockIdx.z;
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = B.lo[d] - (d < g.dim ? 1 : 0);
        hi[d] = B.hi[d] + (d < g.dim ? 1 : 0);
    }
    long long idx = grid_point(g);
    int i, j, k;
    if (!region_point(idx, lo, hi, i, j, k)) return;
    // every direction is computed before any store (loads and stores share
    // one in-order counter, so a load issued after a store waits for it)
    double r0[3] = {0.0, 0.0, 0.0}, r1[3] = {0.0, 0.0, 0.0};
    double f0[3] = {0.0, 0.0, 0.0}, f1[3] = {0.0, 0.0, 0.0};
    vp_ppm_dir<PPM, 0>(B, g, A, n, i, j, k, r0[0], r1[0], f0[0], f1[0]);
    vp_ppm_dir<PPM, 1>(B, g, A, n, i, j, k, r0[1], r1[1], f0[1], f1[1]);
    if (g.dim == 3) vp_ppm_dir<PPM, 2>(B, g, A, n, i, j, k, r0[2], r1[2], f0[2], f1[2]);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (d >= g.dim) break;
        if (PPM == 0) {
            B.Ip[d](i, j, k, n) = r0[d];
            continue;
        }
        B.Ip[n](i, j, k, d) = r0[d];
        B.Im[n](i, j, k, d) = r1[d];
        if (g.ppm_trace_forces == 1) {
            B.Ipf[n](i, j, k, d) = f0[d];
            B.Imf[n](i, j, k, d) = f1[d];
        }
    }
}

// interface states ul/ur of velocity component n on D-faces (VelPredInterface)
template <int D>
__device__ __forceinline__ void vp_iface(const VpBox& B, const KGeom& g, int n, int i, int j,
                                         int k, double& l, double& r) {
    const int im = i - e(D, 0), jm = j - e(D, 1), km = k - e(D, 2);
    if (g.ppm_type == 0) {
        const double dt2 = 0.5 * g.dt;
        const double hd = g.dx[D];
        const DA& SL = B.Ip[D];
        if (g.dim == 3) {
            double maxu = 0.5 - dt2 * dmax(0.0, B.uf(im, jm, km, D)) / hd;
            double minu = 0.5 + dt2 * dmin(0.0, B.uf(i, j, k, D)) / hd;
            l = B.ut(im, jm, km, n) + maxu * SL(im, jm, km, n);
            r = B.ut(i, j, k, n) - minu * SL(i, j, k, n);
        } else {
            double maxu = dmax(0.0, B.uf(im, jm, km, D));
            double minu = dmin(0.0, B.uf(i, j, k, D));
            l = B.ut(im, jm, km, n) + (0.5 - (dt2 / hd) * maxu) * SL(im, jm, km, n);
            r = B.ut(i, j, k, n) - (0.5 + (dt2 / hd) * minu) * SL(i, j, k, n);
        }
    } else {
        l = B.Ip[n](im, jm, km, D);
        r = B.Im[n](i, j, k, D);
    }
    const int pos = D == 0 ? i : (D == 1 ? j : k);
    if (pos == g.domlo[D]) {
        switch (g.phys_lo[D]) {
            case MX_PHYS_INFLOW:
                l = B.ut(im, jm, km, n);
                r = l;
                break;
            case MX_PHYS_SLIPWALL:
            case MX_PHYS_SYMMETRY:
                if (n == D) {
                    l = 0.0;
                    r = 0.0;
                } else {
                    l = r;
                }
                break;
            case MX_PHYS_NOSLIPWALL:
                l = 0.0;
                r = 0.0;
                break;
            case MX_PHYS_OUTFLOW:
                if (g.dim == 2 && D == 0) {
                    r = l;  // 2-D x: urx(n) = ulx(n) after clamping urx(0)
                } else {
                    if (n == D) r = dmin(r, 0.0);
                    l = r;
                }
                break;
            default:
                break;
        }
    } else if (pos == g.domhi[D] + 1) {
        switch (g.phys_hi[D]) {
            case MX_PHYS_INFLOW:
                l = B.ut(i, j, k, n);
                r = l;
                break;
            case MX_PHYS_SLIPWALL:
            case MX_PHYS_SYMMETRY:
                if (n == D) {
                    l = 0.0;
                    r = 0.0;
                } else {
                    r = l;
                }
                break;
            case MX_PHYS_NOSLIPWALL:
                l = 0.0;
                r = 0.0;
                break;
            case MX_PHYS_OUTFLOW:
                if (n == D) l = dmax(l, 0.0);
                r = l;
                break;
            default:
                break;
        }
    }
}

// upwind interface state uimh (VelPredInterface, MaestroVelPred.cpp:887-896)
template <int D>
__device__ __forceinline__ double vp_uimh(const VpBox& B, const KGeom& g, int n, int i, int j,
                                          int k) {
    double l, r;
    vp_iface<D>(B, g, n, i, j, k, l, r);
    const double t = B.tr[D](i, j, k);
    double res = t > 0.0 ? l : r;
    res = dabs(t) < g.rel_eps ? 0.5 * (l + r) : res;
    return res;
}

// transverse state on D-faces of component O corrected by direction T
template <int D, int T>
__device__ __forceinline__ double vp_trans_value(const VpBox& B, const KGeom& g, int i, int j,
                                                int k) {
    constexpr int O = 3 - D - T;
    const int im = i - e(D, 0), jm = j - e(D, 1), km = k - e(D, 2);
    const double dt6 = g.dt / 6.0;
    const double ht = g.dx[T];
    const DA& trt = B.tr[T];
    double l, r;
    vp_iface<D>(B, g, O, i, j, k, l, r);
    l = l - (dt6 / ht) * (trt(im + e(T, 0), jm + e(T, 1), km + e(T, 2)) + trt(im, jm, km)) *
                (vp_uimh<T>(B, g, O, im + e(T, 0), jm + e(T, 1), km + e(T, 2)) -
                 vp_uimh<T>(B, g, O, im, jm, km));
    r = r - (dt6 / ht) * (trt(i + e(T, 0), j + e(T, 1), k + e(T, 2)) + trt(i, j, k)) *
                (vp_uimh<T>(B, g, O, i + e(T, 0), j + e(T, 1), k + e(T, 2)) -
                 vp_uimh<T>(B, g, O, i, j, k));
    const int pos = D == 0 ? i : (D == 1 ? j : k);
    if (pos == g.domlo[D]) {
        switch (g.phys_lo[D]) {
            case MX_PHYS_INFLOW:
                l = B.ut(im, jm, km, O);
                r = l;
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
    } else if (pos == g.domhi[D] + 1) {
        switch (g.phys_hi[D]) {
            case MX_PHYS_INFLOW:
                l = B.ut(i, j, k, O);
                r = l;
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
    const double tv = B.tr[D](i, j, k);
    double res = tv > 0.0 ? l : r;
    res = dabs(tv) < g.rel_eps ? 0.5 * (l + r) : res;
    return res;
}

template <int D, int T>
__device__ __forceinline__ void vp_trans(const VpBox& B, const KGeom& g, int i, int j, int k) {
    B.T[pair_index(D, T)](i, j, k) = vp_trans_value<D, T>(B, g, i, j, k);
}

// is (i,j,k) in the region of pair (D, T): the box grown by one in O = 3-D-T
// and by one face on the high side of D
template <int D, int T>
__device__ __forceinline__ bool in_pair(const VpBox& B, int i, int j, int k) {
    constexpr int O = 3 - D - T;
    const int p[3] = {i, j, k};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int lo = B.lo[d] - (d == O ? 1 : 0);
        const int hi = B.hi[d] + (d == O ? 1 : 0) + (d == D ? 1 : 0);
        if (p[d] < lo || p[d] > hi) return false;
    }
    return true;
}

// all six transverse states of the points of one cell in one thread (3-D):
// every load before the stores
__global__ void __launch_bounds__(256) k_vp_trans_all(const VpBox* __restrict__ boxes, KGeom g) {
    const VpBox& B = boxes[blockIdx.y];
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = B.lo[d] - 1;
        hi[d] = B.hi[d] + 1;
    }
    long long idx = grid_point(g);
    int i, j, k;
    if (!region_point(idx, lo, hi, i, j, k)) return;
    const bool in0 = in_pair<0, 1>(B, i, j, k), in1 = in_pair<0, 2>(B, i, j, k);
    const bool in2 = in_pair<1, 0>(B, i, j, k), in3 = in_pair<1, 2>(B, i, j, k);
    const bool in4 = in_pair<2, 0>(B, i, j, k), in5 = in_pair<2, 1>(B, i, j, k);
    double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0, v4 = 0.0, v5 = 0.0;
    if (in0) v0 = vp_trans_value<0, 1>(B, g, i, j, k);
    if (in1) v1 = vp_trans_value<0, 2>(B, g, i, j, k);
    if (in2) v2 = vp_trans_value<1, 0>(B, g, i, j, k);
    if (in3) v3 = vp_trans_value<1, 2>(B, g, i, j, k);
    if (in4) v4 = vp_trans_value<2, 0>(B, g, i, j, k);
    if (in5) v5 = vp_trans_value<2, 1>(B, g, i, j, k);
    if (in0) B.T[0](i, j, k) = v0;
    if (in1) B.T[1](i, j, k) = v1;
    if (in2) B.T[2](i, j, k) = v2;
    if (in3) B.T[3](i, j, k) = v3;
    if (in4) B.T[4](i, j, k) = v4;
    if (in5) B.T[5](i, j, k) = v5;
}

__global__ void __launch_bounds__(256) k_vp_trans(const VpBox* __restrict__ boxes, KGeom g) {
    const VpBox& B = boxes[blockIdx.y];
    const int p = blockIdx.z;  // pair index
    const int D = p / 2;
    const int q = p % 2;
    const int T = (q < D) ? q : q + 1;
    const int O = 3 - D - T;
    int lo[3] = {B.lo[0], B.lo[1], B.lo[2]};
    int hi[3] = {B.hi[0], B.hi[1], B.hi[2]};
    lo[O] -= 1;
    hi[O] += 1;
    hi[D] += 1;
    long long idx = grid_point(g);
    int i, j, k;
    if (!region_point(idx, lo, hi, i, j, k)) return;
    switch (p) {
        case 0: vp_trans<0, 1>(B, g, i, j, k); break;
        case 1: vp_trans<0, 2>(B, g, i, j, k); break;
        case 2: vp_trans<1, 0>(B, g, i, j, k); break;
        case 3: vp_trans<1, 2>(B, g, i, j, k); break;
        case 4: vp_trans<2, 0>(B, g, i, j, k); break;
        default: vp_trans<2, 1>(B, g, i, j, k); break;
    }
}

// one transverse correction term of VelPredVelocities for direction T
template <int D, int T>
__device__ __forceinline__ void vp_vel_term(const VpBox& B, const KGeom& g, int i, int j, int k,
                                            double& vl, double& vr) {
    const int im = i - e(D, 0), jm = j - e(D, 1), km = k - e(D, 2);
    const double dt4 = 0.25 * g.dt;
    const double ht = g.dx[T];
    const DA& trt = B.tr[T];
    double aL, bL, aR, bR;
    if (g.dim == 3) {
        const DA& Tt = B.T[pair_index(T, 3 - D - T)];
        aL = Tt(im + e(T, 0), jm + e(T, 1), km + e(T, 2));
        bL = Tt(im, jm, km);
        aR = Tt(i + e(T, 0), j + e(T, 1), k + e(T, 2));
        bR = Tt(i, j, k);
    } else {
        aL = vp_uimh<T>(B, g, D, im + e(T, 0), jm + e(T, 1), km + e(T, 2));
        bL = vp_uimh<T>(B, g, D, im, jm, km);
        aR = vp_uimh<T>(B, g, D, i + e(T, 0), j + e(T, 1), k + e(T, 2));
        bR = vp_uimh<T>(B, g, D, i, j, k);
    }
    vl = vl - (dt4 / ht) * (trt(im + e(T, 0), jm + e(T, 1), km + e(T, 2)) + trt(im, jm, km)) *
                  (aL - bL);
    vr = vr - (dt4 / ht) * (trt(i + e(T, 0), j + e(T, 1), k + e(T, 2)) + trt(i, j, k)) * (aR - bR);
}

template <int D>
__device__ __forceinline__ void vp_vel(const VpBox& B, const KGeom& g, int i, int j, int k) {
    const int im = i - e(D, 0), jm = j - e(D, 1), km = k - e(D, 2);
    const double dt2 = 0.5 * g.dt;
    double fl, fr;
    if (g.ppm_trace_forces == 0) {
        fl = B.fo(im, jm, km, D);
        fr = B.fo(i, j, k, D);
    } else {
        fl = B.Ipf[D](im, jm, km, D);
        fr = B.Imf[D](i, j, k, D);
    }
    double vl, vr;
    vp_iface<D>(B, g, D, i, j, k, vl, vr);
    if (D == 0) {
        vp_vel_term<0, 1>(B, g, i, j, k, vl, vr);
        if (g.dim == 3) vp_vel_term<0, 2>(B, g, i, j, k, vl, vr);
    } else if (D == 1) {
        vp_vel_term<1, 0>(B, g, i, j, k, vl, vr);
        if (g.dim == 3) vp_vel_term<1, 2>(B, g, i, j, k, vl, vr);
    } else {
        vp_vel_term<2, 0>(B, g, i, j, k, vl, vr);
        vp_vel_term<2, 1>(B, g, i, j, k, vl, vr);
    }
    const double umacl = vl + dt2 * fl;
    const double umacr = vr + dt2 * fr;

    double res;
    const bool sph = g.dim == 3 && g.spherical;
    if (sph || D == g.dim - 1) {
        const double w = sph ? B.w0m[D](i, j, k) : B.w0c(i, j, k, g.dim - 1);
        bool test = (umacl + w <= 0.0 && umacr + w >= 0.0) ||
                    (dabs(umacl + umacr + 2.0 * w) < g.rel_eps);
        res = 0.5 * (umacl + umacr) + w > 0.0 ? umacl : umacr;
        res = test ? 0.0 : res;
    } else {
        bool test = (umacl <= 0.0 && umacr >= 0.0) || (dabs(umacl + umacr) < g.rel_eps);
        res = 0.5 * (umacl + umacr) > 0.0 ? umacl : umacr;
        res = test ? 0.0 : res;
    }
    const int pos = D == 0 ? i : (D == 1 ? j : k);
    if (pos == g.domlo[D]) {
        switch (g.phys_lo[D]) {
            case MX_PHYS_INFLOW: res = B.ut(im, jm, km, D); break;
            case MX_PHYS_SLIPWALL:
            case MX_PHYS_SYMMETRY:
            case MX_PHYS_NOSLIPWALL: res = 0.0; break;
            case MX_PHYS_OUTFLOW: res = dmin(umacr, 0.0); break;
            default: break;
        }
    } else if (pos == g.domhi[D] + 1) {
        switch (g.phys_hi[D]) {
            case MX_PHYS_INFLOW: res = B.ut(i, j, k, D); break;
            case MX_PHYS_SLIPWALL:
            case MX_PHYS_SYMMETRY:
            case MX_PHYS_NOSLIPWALL: res = 0.0; break;
            case MX_PHYS_OUTFLOW: res = dmax(umacl, 0.0); break;
            default: break;
        }
    }
    B.um[D](i, j, k) = res;
}

__global__ void __launch_bounds__(256) k_vp_vel(const VpBox* __restrict__ boxes, KGeom g) {
    const VpBox& B = boxes[blockIdx.y];
    const int d = blockIdx.z;
    int lo[3] = {B.lo[0], B.lo[1], B.lo[2]};
    int hi[3] = {B.hi[0], B.hi[1], B.hi[2]};
    hi[d] += 1;
    long long idx = grid_point(g);
    int i, j, k;
    if (!region_point(idx, lo, hi, i, j, k)) return;
    if (d == 0)
        vp_vel<0>(B, g, i, j, k);
    else if (d == 1)
        vp_vel<1>(B, g, i, j, k);
    else
        vp_vel<2>(B, g, i, j, k);
}

// MakeUtrans inside the pass: the left and right states of D-face (i,j,k)
// are VelPred's traced states of the normal component (PPM: Ip of the cell
// below, Im of the cell above; ppm_type 0: the slopes), which are the same
// numbers MakeUtrans's own reconstruction gives (same stencil, limiter,
// boundary stencils and tracing velocity ufull)
template <int PPM, int D>
__device__ __forceinline__ void utrans_pp_face(const VpBox& B, const KGeom& g, int i, int j,
                                               int k) {
    const int im = i - e(D, 0), jm = j - e(D, 1), km = k - e(D, 2);
    double ul, ur;
    if (PPM == 0) {
        const double dt2 = 0.5 * g.dt;
        const double hd = g.dx[D];
        const double slL = B.Ip[D](im, jm, km, D);
        const double slR = B.Ip[D](i, j, k, D);
        ul = B.ut(im, jm, km, D) + (0.5 - (dt2 / hd) * dmax(0.0, B.uf(im, jm, km, D))) * slL;
        ur = B.ut(i, j, k, D) - (0.5 + (dt2 / hd) * dmin(0.0, B.uf(i, j, k, D))) * slR;
    } else {
        ul = B.Ip[D](im, jm, km, D);
        ur = B.Im[D](i, j, k, D);
    }
    utrans_riemann<D>(B, g, i, j, k, ul, ur);
}

template <int PPM>
__global__ void __launch_bounds__(256) k_utrans_pp(const VpBox* __restrict__ boxes, KGeom g) {
    const VpBox& B = boxes[blockIdx.y];
    const int d = blockIdx.z;
    int lo[3] = {B.lo[0], B.lo[1], B.lo[2]};
    int hi[3] = {B.hi[0], B.hi[1], B.hi[2]};
    hi[d] += 1;
    long long idx = grid_point(g);
    int i, j, k;
    if (!region_point(idx, lo, hi, i, j, k)) return;
    if (d == 0)
        utrans_pp_face<PPM, 0>(B, g, i, j, k);
    else if (d == 1)
        utrans_pp_face<PPM, 1>(B, g, i, j, k);
    else
        utrans_pp_face<PPM, 2>(B, g, i, j, k);
}

}  // namespace

namespace mxh {

int check_phys(const mx_adv_ctx* c) {
    for (int d = 0; d < c->p.dim; ++d) {
        if (c->p.phys_bc_lo[d] < MX_PHYS_INTERIOR || c->p.phys_bc_lo[d] > MX_PHYS_NOSLIPWALL)
            return MX_ERR_BC;
        if (c->p.phys_bc_hi[d] < MX_PHYS_INTERIOR || c->p.phys_bc_hi[d] > MX_PHYS_NOSLIPWALL)
            return MX_ERR_BC;
    }
    return MX_OK;
}

int make_utrans(mx_adv_ctx* c, int lev, const mx_multifab* utilde, const mx_multifab* ufull,
                const mx_face_multifab* utrans, const mx_face_multifab* w0mac,
                const mx_multifab* w0_cart, bool fill) {
    const int dim = c->p.dim;
    const auto& grids = c->grids[lev];
    const bool sph = dim == 3 && c->p.spherical;
    std::vector<UtBox> boxes;
    long long maxpts = 0;
    for (size_t b = 0; b < grids.size(); ++b) {
        UtBox B;
        HBox g = hbox(grids[b]);
        for (int q = 0; q < 3; ++q) {
            B.lo[q] = g.lo[q];
            B.hi[q] = g.hi[q];
        }
        B.ut = to_da(utilde->fab[b]);
        B.uf = to_da(ufull->fab[b]);
        B.w0c = (!sph && w0_cart) ? to_da(w0_cart->fab[b]) : null_da();
        for (int d = 0; d < 3; ++d) {
            B.tr[d] = d < dim ? to_da(utrans->dir[d].fab[b]) : null_da();
            B.w0m[d] = (sph && d < dim) ? to_da(w0mac->dir[d].fab[b]) : null_da();
        }
        // the grid covers the largest face region of any direction
        for (int d = 0; d < dim; ++d) maxpts = std::max(maxpts, npts(nodal(g, d)));
        boxes.push_back(B);
    }
    // the line form: 3-D, ppm_type 1 and 2; blocks for the larger of the
    // x-face rows and the y/z column chunks of any box
    const int nch = c->utrans_chunks;
    const bool line = c->utrans_line && dim == 3 && c->p.ppm_type != 0;
    long long lblocks = 1;
    if (line)
        for (const UtBox& B : boxes) {
            const long long n0 = B.hi[0] - B.lo[0] + 1, n1 = B.hi[1] - B.lo[1] + 1,
                            n2 = B.hi[2] - B.lo[2] + 1;
            const long long rows = (n0 + 2) * n1 * n2;
            if (rows >= 0x7fffffffLL || n0 * std::max(n1, n2) * nch >= 0x7fffffffLL)
                return MX_ERR_UNSUPPORTED;
            // wavefronts of 63 new cells each, 4 per block
            const long long waves = (rows - 1 + 62) / 63;
            lblocks = std::max(lblocks, (waves + UT_NT / 64 - 1) / (UT_NT / 64));
            lblocks = std::max(lblocks, (n0 * std::max(n1, n2) * nch + UT_NT - 1) / UT_NT);
        }
    UtArgs A;
    for (int n = 0; n < 3; ++n)
        for (int d = 0; d < 3; ++d) {
            A.bclo[n][d] = c->bcs_u[n].lo[d];
            A.bchi[n][d] = c->bcs_u[n].hi[d];
        }
    Table tab(c);
    size_t off = tab.add(boxes);
    int rc = tab.upload();
    if (rc) return rc;
    KGeom kg = kgeom(c, lev);
    const int block = 256;
    const UtBox* dp = tab.dev<UtBox>(off);
    dim3 gr = grid_for(maxpts, (int)boxes.size(), dim, block);
    const dim3 gl((unsigned)lblocks, (unsigned)boxes.size(), 3);
    if (line && c->p.ppm_type == 1)
        MX_LAUNCH(c, "k_utrans", k_utrans_line<1>, gl, dim3(UT_NT), dp, kg, A, nch);
    else if (line)
        MX_LAUNCH(c, "k_utrans", k_utrans_line<2>, gl, dim3(UT_NT), dp, kg, A, nch);
    else if (c->p.ppm_type == 0)
        MX_LAUNCH(c, "k_utrans", k_utrans<0>,  gr, dim3(block), dp, kg, A);
    else if (c->p.ppm_type == 1)
        MX_LAUNCH(c, "k_utrans", k_utrans<1>,  gr, dim3(block), dp, kg, A);
    else
        MX_LAUNCH(c, "k_utrans", k_utrans<2>,  gr, dim3(block), dp, kg, A);

    return fill ? fill_face_ghosts(c, lev, utrans) : MX_OK;
}

int velpred_setup(mx_adv_ctx* c, int lev, const mx_multifab* utilde, const mx_multifab* ufull,
                  const mx_face_multifab* utrans, const mx_face_multifab* umac,
                  const mx_face_multifab* w0mac, const mx_multifab* force,
                  const mx_multifab* w0_cart, VpPlan& plan, bool reserve) {
    const int dim = c->p.dim;
    const auto& grids = c->grids[lev];
    const bool sph = dim == 3 && c->p.spherical;
    const int ppm = c->p.ppm_type;
    const bool tf = c->p.ppm_trace_forces == 1 && ppm != 0;
    if (dim == 2 && c->p.ppm_trace_forces == 1) return MX_ERR_UNSUPPORTED;
    int rc = check_phys(c);
    if (rc) return rc;

    auto grown1 = [&](HBox b) { return grow(b, dim, 1); };
    // reserve = false: the caller reserved the workspace of several levels
    // and each level takes its temporaries in turn
    if (reserve) {
        rc = ws_reserve(c, velpred_ws_bytes(c, lev));
        if (rc) return rc;
    }

    std::vector<DA> Ip[3], Im[3], Ipf[3], Imf[3], T[6];
    for (int n = 0; n < dim; ++n) {
        Ip[n] = tmp_fabs(c, grids, dim, grown1);
        Im[n] = tmp_fabs(c, grids, dim, grown1);
        if (tf) {
            Ipf[n] = tmp_fabs(c, grids, dim, grown1);
            Imf[n] = tmp_fabs(c, grids, dim, grown1);
        }
    }
    if (dim == 3)
        for (int p = 0; p < 6; ++p) T[p] = tmp_fabs(c, grids, 1, grown1);

    std::vector<VpBox> boxes;
    long long maxpts = 0;
    for (size_t b = 0; b < grids.size(); ++b) {
        VpBox B;
        HBox g = hbox(grids[b]);
        for (int q = 0; q < 3; ++q) {
            B.lo[q] = g.lo[q];
            B.hi[q] = g.hi[q];
        }
        B.ut = to_da(utilde->fab[b]);
        B.uf = to_da(ufull->fab[b]);
        B.fo = to_da(force->fab[b]);
        B.w0c = (!sph && w0_cart) ? to_da(w0_cart->fab[b]) : null_da();
        for (int d = 0; d < 3; ++d) {
            B.tr[d] = d < dim ? to_da(utrans->dir[d].fab[b]) : null_da();
            B.um[d] = d < dim ? to_da(umac->dir[d].fab[b]) : null_da();
            B.w0m[d] = (sph && d < dim) ? to_da(w0mac->dir[d].fab[b]) : null_da();
            B.Ip[d] = d < dim ? Ip[d][b] : null_da();
            B.Im[d] = d < dim ? Im[d][b] : null_da();
            B.Ipf[d] = (tf && d < dim) ? Ipf[d][b] : null_da();
            B.Imf[d] = (tf && d < dim) ? Imf[d][b] : null_da();
        }
        for (int p = 0; p < 6; ++p) B.T[p] = dim == 3 ? T[p][b] : null_da();
        maxpts = std::max(maxpts, npts(grown1(g)));
        boxes.push_back(B);
    }
    VpArgs A;
    for (int n = 0; n < 3; ++n)
        for (int d = 0; d < 3; ++d) {
            A.bclo[n][d] = c->bcs_u[n].lo[d];
            A.bchi[n][d] = c->bcs_u[n].hi[d];
        }
    Table tab(c);
    size_t off = tab.add(boxes);
    rc = tab.upload();
    if (rc) return rc;
    const VpBox* dp = tab.dev<VpBox>(off);
    KGeom kg = kgeom(c, lev);
    const int block = 256;
    const int nb = (int)boxes.size();

    dim3 gr = grid_for_g(kg, maxpts, nb, dim, block);
    if (c->vp_ppm_all && dim == 3 && ppm != 0 && c->p.ppm_trace_forces == 0) {
        const dim3 g1 = grid_for_g(kg, maxpts, nb, 1, block);
        if (ppm == 1)
            MX_LAUNCH(c, "k_vp_ppm", k_vp_ppm_all<1>, g1, dim3(block), dp, kg, A);
        else
            MX_LAUNCH(c, "k_vp_ppm", k_vp_ppm_all<2>, g1, dim3(block), dp, kg, A);
    } else if (ppm == 0)
        MX_LAUNCH(c, "k_vp_ppm", k_vp_ppm<0>,  gr, dim3(block), dp, kg, A);
    else if (ppm == 1)
        MX_LAUNCH(c, "k_vp_ppm", k_vp_ppm<1>,  gr, dim3(block), dp, kg, A);
    else
        MX_LAUNCH(c, "k_vp_ppm", k_vp_ppm<2>,  gr, dim3(block), dp, kg, A);
    plan.dp = dp;
    plan.nb = nb;
    plan.maxpts = maxpts;
    plan.lev = lev;
    plan.utrans = utrans;
    return MX_OK;
}

// VelPred's temporaries on one level: Ip/Im (or slopes) per component,
// traced forces, transverse states, each over the boxes grown by one
size_t velpred_ws_bytes(const mx_adv_ctx* c, int lev) {
    const int dim = c->p.dim;
    const auto& grids = c->grids[lev];
    const bool tf = c->p.ppm_trace_forces == 1 && c->p.ppm_type != 0;
    auto grown1 = [&](HBox b) { return grow(b, dim, 1); };
    size_t need = 2 * dim * tmp_bytes(grids, dim, grown1);
    if (tf) need += 2 * dim * tmp_bytes(grids, dim, grown1);
    if (dim == 3) need += 6 * tmp_bytes(grids, 1, grown1);
    return need;
}

// MakeUtrans from the traced states of velpred_setup, then the face ghosts
// (fill = false: the caller synchronises the levels instead)
int utrans_from_plan(mx_adv_ctx* c, const VpPlan& plan, bool fill) {
    const KGeom kg = kgeom(c, plan.lev);
    const int block = 256;
    const dim3 gr = grid_for_g(kg, plan.maxpts, plan.nb, c->p.dim, block);
    if (c->p.ppm_type == 0)
        MX_LAUNCH(c, "k_utrans", k_utrans_pp<0>, gr, dim3(block), (const VpBox*)plan.dp, kg);
    else
        MX_LAUNCH(c, "k_utrans", k_utrans_pp<1>, gr, dim3(block), (const VpBox*)plan.dp, kg);
    return fill ? fill_face_ghosts(c, plan.lev, plan.utrans) : MX_OK;
}

// VelPred after its PPM kernel: transverse states and MAC velocities
int velpred_finish(mx_adv_ctx* c, const VpPlan& plan) {
    const KGeom kg = kgeom(c, plan.lev);
    const int block = 256;
    if (c->p.dim == 3 && c->vp_trans_all)
        MX_LAUNCH(c, "k_vp_trans", k_vp_trans_all, grid_for_g(kg, plan.maxpts, plan.nb, 1, block),
                  dim3(block), (const VpBox*)plan.dp, kg);
    else if (c->p.dim == 3)
        MX_LAUNCH(c, "k_vp_trans", k_vp_trans, grid_for_g(kg, plan.maxpts, plan.nb, 6, block),
                  dim3(block), (const VpBox*)plan.dp, kg);
    MX_LAUNCH(c, "k_vp_vel", k_vp_vel, grid_for_g(kg, plan.maxpts, plan.nb, c->p.dim, block),
              dim3(block), (const VpBox*)plan.dp, kg);
    // AverageDownFaces(umac) is a no-op on a single level
    return MX_OK;
}

int velpred(mx_adv_ctx* c, int lev, const mx_multifab* utilde, const mx_multifab* ufull,
            const mx_face_multifab* utrans, const mx_face_multifab* umac,
            const mx_face_multifab* w0mac, const mx_multifab* force, const mx_multifab* w0_cart) {
    if (velpred_fused_ok(c)) {
        std::vector<VpLevel> levs(1);
        levs[0] = {lev, utilde, ufull, utrans, umac, w0mac, force, w0_cart};
        return velpred_fused(c, levs);
    }
    VpPlan plan;
    int rc = velpred_setup(c, lev, utilde, ufull, utrans, umac, w0mac, force, w0_cart, plan);
    if (rc) return rc;
    return velpred_finish(c, plan);
}

}  // namespace mxh

extern "C" {

int mx_adv_make_utrans(mx_adv_ctx* c, int lev, const mx_multifab* utilde,
                       const mx_multifab* ufull, mx_face_multifab* utrans,
                       const mx_face_multifab* w0mac, const mx_multifab* w0_cart) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    const int dim = c->p.dim;
    const int g = c->p.ppm_type == 2 ? 4 : 3;
    if ((rc = check_mf(c, lev, utilde, dim, g, -1))) return rc;
    if ((rc = check_mf(c, lev, ufull, dim, 1, -1))) return rc;
    if ((rc = check_face(c, lev, utrans, 1, 1))) return rc;
    if (dim == 3 && c->p.spherical) {
        if ((rc = check_face(c, lev, w0mac, 1, 0))) return rc;
    } else {
        if ((rc = check_mf(c, lev, w0_cart, dim, 1, -1))) return rc;
    }
    return make_utrans(c, lev, utilde, ufull, utrans, w0mac, w0_cart);
}

int mx_adv_velpred(mx_adv_ctx* c, int lev, const mx_multifab* utilde, const mx_multifab* ufull,
                   const mx_face_multifab* utrans, mx_face_multifab* umac,
                   const mx_face_multifab* w0mac, const mx_multifab* force,
                   const mx_multifab* w0_cart) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    const int dim = c->p.dim;
    const int g = c->p.ppm_type == 2 ? 4 : 3;
    if ((rc = check_mf(c, lev, utilde, dim, g, -1))) return rc;
    if ((rc = check_mf(c, lev, ufull, dim, 1, -1))) return rc;
    if ((rc = check_face(c, lev, utrans, 1, 1))) return rc;
    if ((rc = check_face(c, lev, umac, 1, 0))) return rc;
    const int fg = (c->p.ppm_trace_forces == 1) ? g : 1;
    if ((rc = check_mf(c, lev, force, dim, fg, -1))) return rc;
    if (dim == 3 && c->p.spherical) {
        if ((rc = check_face(c, lev, w0mac, 1, 0))) return rc;
    } else {
        if ((rc = check_mf(c, lev, w0_cart, dim, 1, -1))) return rc;
    }
    return velpred(c, lev, utilde, ufull, utrans, umac, w0mac, force, w0_cart);
}

}  // extern "C"
