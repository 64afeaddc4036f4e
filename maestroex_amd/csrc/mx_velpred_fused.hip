// mx_velpred_fused.hip -- fused 3-D VelPred: PPM of the three velocity
// components, the interface, transverse and final states and the Riemann
// solve of the MAC velocities in one z-marching kernel
// (MaestroVelPred.cpp:5-292 with VelPredInterface :735-1128,
// VelPredTransverse :1130-1595 and VelPredVelocities :1597-1915).
//
// A block owns a TX x TY tile of cell columns of one box, grown by one cell
// in x and y (one thread per column), and marches up a chunk [z0, z1] of the
// box's planes, one cell plane q per step (plus q = z0 - 1 and q = z1 + 1).
// Nothing goes through HBM between the stages: the reference's sixteen
// temporaries (Ip/Im per component, the interface and upwind states, the six
// transverse states) live in LDS and registers.
//
// Names: iface(D, n) are the left/right interface states of velocity
// component n on D-faces, uimh(T, n) their upwind state by utrans; the
// transverse state pair (D, T) is component O = 3 - D - T on D-faces
// corrected by direction T (uimhyz = (1, 2), uimhzy = (2, 1), vimhxz =
// (0, 2), vimhzx = (2, 0), wimhxy = (0, 1), wimhyx = (1, 0)); umac on
// D-faces is corrected by the pairs (T, 3 - D - T), T != D.
//
// The reference formulas are regrouped, not changed: a transverse correction
// (dt/6h_T)(utrans_T(+) + utrans_T(-))(uimh_T(+) - uimh_T(-)) depends only on
// the cell it is taken across, so it is evaluated once per cell and read by
// the faces on both sides; the final corrections (dt/4h_T) likewise.  Each is
// the reference's expression in the reference's operand order, so the
// results are bit-identical to the per-point kernels.  PPM boundary stencils
// (ext_dir / hoextrap within three cells of a domain side) use the full
// per-cell routine of mx_device.H; the tracing velocity is ufull at the cell
// (is_umac = 0).
//
// Per step q, phases separated by barriers:
//   A  stage plane q of the three components and of utrans_x, utrans_y;
//      shift the column's z windows; issue the loads of plane q + 1;
//   B1 face values of plane q (x, y, three components); PPM of cell q in z
//      (register windows) and the z-face q interface / upwind states;
//   B2 PPM of plane q in x and y;
//   C  interface and upwind states on the x- and y-faces of plane q;
//   D1 per-cell transverse corrections: x, y of plane q, z of plane q - 1;
//   D2 transverse states wimhxy, wimhyx (plane q), vimhxz, uimhyz (plane
//      q - 1), vimhzx, uimhzy (z-face q, registers);
//   E1 per-cell final corrections: plane q - 1 (x, y faces), plane q (z);
//   E2 umac, vmac on the faces of plane q - 1 and wmac on z-face q.
#include <algorithm>

#include "mx_internal.H"

using namespace mxd;
using namespace mxh;

namespace {

typedef __attribute__((address_space(1))) const double gcd;
typedef __attribute__((address_space(1))) double gd;

__host__ __device__ constexpr int vf_nt(int tx, int ty) {
    return ((tx + 2) * (ty + 2) + 63) / 64 * 64;
}

struct VFBox {
    int lo[3], hi[3];
    DA ut, uf, fo, w0c;
    DA tr[3], um[3], w0m[3];
    double dx[3];
    int domlo[3], domhi[3];
};

struct VFTile {
    int box, tx, ty, tzc;
};

struct VFArgs {
    int bclo[3][3], bchi[3][3];  // bcs_u[n].lo/hi[dir] (BCType)
    int plo[3], phi[3];          // phys_bc lo/hi (PhysBCType)
    int spherical;
    int tz;
};

// VelPredInterface overrides of the (l, r) pair of a component on a face at
// `pos` (3-D; MaestroVelPred.cpp:790-875): `normal` when the component is the
// face's normal velocity, ulo / uhi the cell values below / above the face
__device__ __forceinline__ void vf_iface_bc(int pos, int dlo, int dhi, int bclo, int bchi,
                                            bool normal, double ulo, double uhi, double& l,
                                            double& r) {
    if (pos == dlo) {
        switch (bclo) {
            case MX_PHYS_INFLOW:
                l = ulo;
                r = l;
                break;
            case MX_PHYS_SLIPWALL:
            case MX_PHYS_SYMMETRY:
                if (normal) {
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
                if (normal) r = dmin(r, 0.0);
                l = r;
                break;
            default:
                break;
        }
    } else if (pos == dhi + 1) {
        switch (bchi) {
            case MX_PHYS_INFLOW:
                l = uhi;
                r = l;
                break;
            case MX_PHYS_SLIPWALL:
            case MX_PHYS_SYMMETRY:
                if (normal) {
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
                if (normal) l = dmax(l, 0.0);
                r = l;
                break;
            default:
                break;
        }
    }
}

// VelPredTransverse overrides after the correction (MaestroVelPred.cpp:1183-1230)
__device__ __forceinline__ void vf_trans_bc(int pos, int dlo, int dhi, int bclo, int bchi,
                                            double ulo, double uhi, double& l, double& r) {
    if (pos == dlo) {
        switch (bclo) {
            case MX_PHYS_INFLOW:
                l = ulo;
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
    } else if (pos == dhi + 1) {
        switch (bchi) {
            case MX_PHYS_INFLOW:
                l = uhi;
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
}

// upwind by a transverse velocity, averaging below rel_eps (MaestroVelPred.cpp:887-896)
__device__ __forceinline__ double vf_upwind(double l, double r, double t, double rel_eps) {
    double res = t > 0.0 ? l : r;
    res = dabs(t) < rel_eps ? 0.5 * (l + r) : res;
    return res;
}

// the Riemann solve of VelPredVelocities, with w0 where it acts
// (MaestroVelPred.cpp:1690-1715)
__device__ __forceinline__ double vf_riemann(double umacl, double umacr, bool has_w, double w,
                                             double rel_eps) {
    if (has_w) {
        const bool test = (umacl + w <= 0.0 && umacr + w >= 0.0) ||
                          (dabs(umacl + umacr + 2.0 * w) < rel_eps);
        const double res = 0.5 * (umacl + umacr) + w > 0.0 ? umacl : umacr;
        return test ? 0.0 : res;
    }
    const bool test = (umacl <= 0.0 && umacr >= 0.0) || (dabs(umacl + umacr) < rel_eps);
    const double res = 0.5 * (umacl + umacr) > 0.0 ? umacl : umacr;
    return test ? 0.0 : res;
}

// the overrides after the Riemann solve (MaestroVelPred.cpp:1717-1750)
__device__ __forceinline__ double vf_vel_bc(double res, int pos, int dlo, int dhi, int bclo,
                                            int bchi, double ulo, double uhi, double umacl,
                                            double umacr) {
    if (pos == dlo) {
        switch (bclo) {
            case MX_PHYS_INFLOW: res = ulo; break;
            case MX_PHYS_SLIPWALL:
            case MX_PHYS_SYMMETRY:
            case MX_PHYS_NOSLIPWALL: res = 0.0; break;
            case MX_PHYS_OUTFLOW: res = dmin(umacr, 0.0); break;
            default: break;
        }
    } else if (pos == dhi + 1) {
        switch (bchi) {
            case MX_PHYS_INFLOW: res = uhi; break;
            case MX_PHYS_SLIPWALL:
            case MX_PHYS_SYMMETRY:
            case MX_PHYS_NOSLIPWALL: res = 0.0; break;
            case MX_PHYS_OUTFLOW: res = dmax(umacl, 0.0); break;
            default: break;
        }
    }
    return res;
}

// PPM edge values of a cell from its face values (type 1) or its stencil and
// face values (type 2), with the boundary stencils where they act
template <int PPM, int D>
__device__ __forceinline__ void vf_cell_edges(const double* w9, const double* F, int fst, int pos,
                                              int dlo, int dhi, int bclo, int bchi,
                                              bool check_bc, double& sm, double& sp) {
    if (PPM == 1) {
        sm = F[0];
        sp = F[fst];
        cw_limit(w9[4], sm, sp);
    } else {
        cs_cell(w9, F[-fst], F[0], F[fst], F[2 * fst], 1.e-10, sm, sp);
    }
    if (check_bc && ppm_special<PPM>(pos, dlo, dhi, bclo, bchi))
        ppm_edges<PPM, D>(w9, pos, dlo, dhi, bclo, bchi, sm, sp);
}

// XYF: no x- or y-boundary condition acts on the tile (no PPM boundary
// stencil and no physical-boundary override within the tile's reach)
template <int PPM, int TX, int TY, bool XYF, int WPE>
__global__ void __launch_bounds__(vf_nt(TX, TY)) __attribute__((amdgpu_waves_per_eu(WPE)))
    k_vp_fused(const VFBox* __restrict__ boxes, const VFTile* __restrict__ tiles, KGeom g,
               VFArgs A) {
    constexpr int NT = vf_nt(TX, TY);
    constexpr int G = PPM == 2 ? 3 : 2;
    constexpr int W = TX + 2;
    constexpr int H = TY + 2;
    constexpr int NE = W * H;
    constexpr int SW = W + 2 * G;
    constexpr int SH = H + 2 * G;
    constexpr int UW = W + 1;  // x-faces per row of the utrans_x plane
    constexpr int VH = H + 1;  // y-faces per column of the utrans_y plane
    constexpr int FO = PPM == 2 ? 1 : 0;
    constexpr int FXW = W + 1 + 2 * FO;  // x-face values per row
    constexpr int FYH = H + 1 + 2 * FO;  // y-face values per column
    constexpr int NF = FXW * H > W * FYH ? FXW * H : W * FYH;
    static_assert(NE <= NT, "tile too large for the block");

    __shared__ double S[3][SW * SH];
    __shared__ double UL[3][UW * H];
    __shared__ double VL[3][W * VH];
    __shared__ double FX[3][NF];   // x-face values; later phases reuse it
    __shared__ double FY[3][NF];   // y-face values; later phases reuse it
    __shared__ double P[12][NE];   // Ipx, Imx, Ipy, Imy per component
    __shared__ double SC[3][NE];   // plane q - 1 cell values (inflow overrides)
    __shared__ double FL[2][NE];   // plane q - 1 forces, x and y components

    const VFTile tl = tiles[blockIdx.x];
    const VFBox& B = boxes[tl.box];
    const int bhi0 = B.hi[0], bhi1 = B.hi[1], bhi2 = B.hi[2];
    const int X0 = B.lo[0] + tl.tx * TX;
    const int Y0 = B.lo[1] + tl.ty * TY;
    const int nx = min(TX, bhi0 - X0 + 1);
    const int ny = min(TY, bhi1 - Y0 + 1);
    const int z0 = B.lo[2] + tl.tzc * A.tz;
    const int z1 = min(z0 + A.tz - 1, bhi2);
    const bool last_x = (X0 + nx - 1 == bhi0);
    const bool last_y = (Y0 + ny - 1 == bhi1);
    const bool last_z = (z1 == bhi2);
    const int dlx = B.domlo[0], dhx = B.domhi[0];
    const int dly = B.domlo[1], dhy = B.domhi[1];
    const int dlz = B.domlo[2], dhz = B.domhi[2];
    const int pxl = A.plo[0], pxh = A.phi[0], pyl = A.plo[1], pyh = A.phi[1];
    const int pzl = A.plo[2], pzh = A.phi[2];
    const bool sph = A.spherical != 0;

    const DA utv = B.ut, ufv = B.uf, fov = B.fo;
    const DA trx = B.tr[0], tryv = B.tr[1], trz = B.tr[2];
    const DA umx = B.um[0], umy = B.um[1], umz = B.um[2];
    const int sjs = (int)utv.js, sks = (int)utv.ks, sns = (int)utv.ns;
    const int xjs = (int)trx.js, xks = (int)trx.ks;
    const int yjs = (int)tryv.js, yks = (int)tryv.ks;

    const int tid = threadIdx.x;
    const int a = tid % W;  // local column: cell (X0 + a - 1, Y0 + b - 1)
    const int b = tid / W;
    const bool col = (tid < NE) && (a <= nx + 1) && (b <= ny + 1);
    const int I = X0 + a - 1;
    const int J = Y0 + b - 1;
    const int si = (a + G) + (b + G) * SW;
    const int ui = a + b * UW;
    const int vi = a + b * W;

    const double dt = g.dt;
    const double rel_eps = g.rel_eps;
    const double dt2 = 0.5 * dt;
    const double dt4 = 0.25 * dt;
    const double dt6 = dt / 6.0;
    const double hx = B.dx[0], hy = B.dx[1], hz = B.dx[2];
    const double t6x = dt6 / hx, t6y = dt6 / hy, t6z = dt6 / hz;
    const double t4x = dt4 / hx, t4y = dt4 / hy, t4z = dt4 / hz;

    // plane pointers (block-uniform) at plane z0 - 1
    gcd* sPl = (gcd*)utv.p + utv.off(X0 - 1 - G, Y0 - 1 - G, z0 - 1);
    gcd* xPl = (gcd*)trx.p + trx.off(X0 - 1, Y0 - 1, z0 - 1);
    gcd* yPl = (gcd*)tryv.p + tryv.off(X0 - 1, Y0 - 1, z0 - 1);
    // column pointers
    gcd* sCol = (gcd*)utv.p;
    gcd* ufCol = (gcd*)ufv.p;
    gcd* foCol = (gcd*)fov.p;
    gcd* zCol = (gcd*)trz.p;
    gd* oxCol = (gd*)umx.p;
    gd* oyCol = (gd*)umy.p;
    gd* ozCol = (gd*)umz.p;
    if (col) {
        sCol += utv.off(I, J, z0 - 1 + G);  // next z-window load: s(q + G)
        ufCol += ufv.off(I, J, z0 - 1);
        foCol += fov.off(I, J, z0 - 1);
        zCol += trz.off(I, J, z0 - 1);
        oxCol += umx.off(I, J, z0 - 1);
        oyCol += umy.off(I, J, z0 - 1);
        ozCol += umz.off(I, J, z0 - 1);
    }
    const int ufks = (int)ufv.ks, ufns = (int)ufv.ns;
    const int foks = (int)fov.ks, fons = (int)fov.ns;
    const int zks = (int)trz.ks;
    const int oxks = (int)umx.ks, oyks = (int)umy.ks, ozks = (int)umz.ks;

    // column state
    double v[3][9];
#pragma unroll
    for (int n = 0; n < 3; ++n)
#pragma unroll
        for (int jj = 0; jj < 9; ++jj) v[n][jj] = 0.0;
    double zf0[3] = {0.0, 0.0, 0.0}, zf1[3] = {0.0, 0.0, 0.0}, zf2[3] = {0.0, 0.0, 0.0};
    double zsl[3] = {0.0, 0.0, 0.0};
    double Ipz_prev[3] = {0.0, 0.0, 0.0};
    double zl[3] = {0.0, 0.0, 0.0}, zr[3] = {0.0, 0.0, 0.0};  // iface(2, n) on z-face q
    double uz_q[2] = {0.0, 0.0}, uz_p[2] = {0.0, 0.0};        // uimh(2, 0), uimh(2, 1)
    double xl0_p = 0.0, xr0_p = 0.0, xl1_p = 0.0, xr1_p = 0.0;  // iface(0, 0/1), plane p
    double yl1_p = 0.0, yr1_p = 0.0, yl0_p = 0.0, yr0_p = 0.0;  // iface(1, 1/0), plane p
    double xl[3] = {0.0, 0.0, 0.0}, xr[3] = {0.0, 0.0, 0.0};
    double yl[3] = {0.0, 0.0, 0.0}, yr[3] = {0.0, 0.0, 0.0};
    double t01_q = 0.0, t01_p = 0.0, t10_q = 0.0, t10_p = 0.0;  // own-column corrections
    double p201_q = 0.0, p201_p = 0.0, p210_q = 0.0, p210_p = 0.0;  // vimhzx, uimhzy
    double ez0_q = 0.0, ez0_p = 0.0, ez1_q = 0.0, ez1_p = 0.0;  // wmac corrections
    double trz_q = 0.0, trz_p = 0.0;
    double uf_q[3] = {0.0, 0.0, 0.0};
    double fo_q[3] = {0.0, 0.0, 0.0}, foz_p = 0.0;
    if (col) {
#pragma unroll
        for (int n = 0; n < 3; ++n)
            for (int jj = 5 - G; jj <= 4 + G; ++jj) v[n][jj] = utv(I, J, z0 - 6 + jj, n);
    }
    int slot_q = 0, slot_p = 2;

    // plane q + 1's loads are issued during step q (the barriers wait for LDS
    // traffic only); each thread owns fixed slots of the staged planes
    constexpr int NS = (SW * SH + NT - 1) / NT;
    constexpr int NU = (UW * H + NT - 1) / NT;
    constexpr int NV = (W * VH + NT - 1) / NT;
    bool okS[NS], okU[NU], okV[NV];
    int offS[NS], offU[NU], offV[NV];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const int idx = tid + k * NT;
        const int sx = idx % SW, sy = idx / SW;
        okS[k] = idx < SW * SH && sx <= nx + 1 + 2 * G && sy <= ny + 1 + 2 * G;
        offS[k] = sx + sy * sjs;
    }
#pragma unroll
    for (int k = 0; k < NU; ++k) {
        const int idx = tid + k * NT;
        const int f = idx % UW, r = idx / UW;
        okU[k] = idx < UW * H && f <= nx + 2 && r <= ny + 1;
        offU[k] = f + r * xjs;
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const int idx = tid + k * NT;
        const int cc = idx % W, r = idx / W;
        okV[k] = idx < W * VH && cc <= nx + 1 && r <= ny + 2;
        offV[k] = cc + r * yjs;
    }
    double pS[3][NS], pU[NU], pV[NV];
    double pcs[3] = {0.0, 0.0, 0.0}, puf[3] = {0.0, 0.0, 0.0}, pfo[3] = {0.0, 0.0, 0.0};
    double pz = 0.0;
#pragma unroll
    for (int k = 0; k < NS; ++k)
#pragma unroll
        for (int n = 0; n < 3; ++n) pS[n][k] = okS[k] ? sPl[offS[k] + n * sns] : 0.0;
#pragma unroll
    for (int k = 0; k < NU; ++k) pU[k] = okU[k] ? xPl[offU[k]] : 0.0;
#pragma unroll
    for (int k = 0; k < NV; ++k) pV[k] = okV[k] ? yPl[offV[k]] : 0.0;
    if (col) {
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            pcs[n] = sCol[n * sns];
            puf[n] = ufCol[n * ufns];
            pfo[n] = foCol[n * fons];
        }
        pz = *zCol;
    }

    for (int q = z0 - 1; q <= z1 + 1; ++q) {
        const bool zface = q >= z0;        // z-face q is in the chunk's range
        const bool pplane = q >= z0 + 1;   // plane p = q - 1 is a chunk plane
        const bool first = q == z0 - 1;
        double* ULq = UL[slot_q];
        double* VLq = VL[slot_q];
        const double* ULp = UL[slot_p];
        const double* VLp = VL[slot_p];

        // ---- A: stage plane q; shift the column windows; issue the loads
        //      of plane q + 1
#pragma unroll
        for (int k = 0; k < NS; ++k)
            if (okS[k]) {
#pragma unroll
                for (int n = 0; n < 3; ++n) S[n][tid + k * NT] = pS[n][k];
            }
#pragma unroll
        for (int k = 0; k < NU; ++k)
            if (okU[k]) ULq[tid + k * NT] = pU[k];
#pragma unroll
        for (int k = 0; k < NV; ++k)
            if (okV[k]) VLq[tid + k * NT] = pV[k];
        double flx = 0.0, fly = 0.0;  // plane q - 1 forces, x and y
        if (col) {
#pragma unroll
            for (int n = 0; n < 3; ++n) {
#pragma unroll
                for (int jj = 0; jj < 8; ++jj) v[n][jj] = v[n][jj + 1];
                v[n][4 + G] = pcs[n];
            }
            flx = fo_q[0];
            fly = fo_q[1];
            foz_p = fo_q[2];
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                uf_q[n] = puf[n];
                fo_q[n] = pfo[n];
            }
            trz_p = trz_q;
            trz_q = pz;
        }
        if (q <= z1) {
            sPl += sks;
            xPl += xks;
            yPl += yks;
#pragma unroll
            for (int k = 0; k < NS; ++k)
                if (okS[k]) {
#pragma unroll
                    for (int n = 0; n < 3; ++n) pS[n][k] = sPl[offS[k] + n * sns];
                }
#pragma unroll
            for (int k = 0; k < NU; ++k)
                if (okU[k]) pU[k] = xPl[offU[k]];
#pragma unroll
            for (int k = 0; k < NV; ++k)
                if (okV[k]) pV[k] = yPl[offV[k]];
            if (col) {
                sCol += sks;
                ufCol += ufks;
                foCol += foks;
                zCol += zks;
#pragma unroll
                for (int n = 0; n < 3; ++n) {
                    pcs[n] = sCol[n * sns];
                    puf[n] = ufCol[n * ufns];
                    pfo[n] = foCol[n * fons];
                }
                pz = *zCol;
            }
        }
        __syncthreads();

        // plane q - 1's cell values and forces, which D2 and E2 read across
        // columns: stored after the barrier above, which no thread passes
        // before every thread is done with the previous plane's E2
        if (col) {
#pragma unroll
            for (int n = 0; n < 3; ++n) SC[n][tid] = v[n][3];
            FL[0][tid] = flx;
            FL[1][tid] = fly;
        }

        // ---- B1: face values of plane q; PPM of cell q in z
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            const double* Sn = S[n];
            for (int idx = tid; idx < FXW * H; idx += NT) {
                const int f = idx % FXW - FO, r = idx / FXW;  // face f: cells f-1 | f
                if (f <= nx + 2 + FO && r <= ny + 1) {
                    const double* sr = Sn + (f + G) + (r + G) * SW;
                    if (PPM == 1) {
                        const double sa2 = sr[-2], sa = sr[-1], sb = sr[0], sb2 = sr[1];
                        FX[n][idx] = face1(sa, sb, vl_slope(sa2, sa, sb), vl_slope(sa, sb, sb2));
                    } else {
                        FX[n][idx] = cs_face(sr[-2], sr[-1], sr[0], sr[1]);
                    }
                }
            }
            for (int idx = tid; idx < W * FYH; idx += NT) {
                const int c = idx % W, f = idx / W - FO;  // face f: cells f-1 | f
                if (c <= nx + 1 && f <= ny + 2 + FO) {
                    const double* sr = Sn + (c + G) + (f + G) * SW;
                    if (PPM == 1) {
                        const double sa2 = sr[-2 * SW], sa = sr[-SW], sb = sr[0], sb2 = sr[SW];
                        FY[n][idx] = face1(sa, sb, vl_slope(sa2, sa, sb), vl_slope(sa, sb, sb2));
                    } else {
                        FY[n][idx] = cs_face(sr[-2 * SW], sr[-SW], sr[0], sr[SW]);
                    }
                }
            }
        }
        if (col) {
            const double sgz = dabs(uf_q[2]) * dt / hz;
            double Ipz[3], Imz[3];
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                double* vn = v[n];
                double sm, sp;
                if (PPM == 1) {
                    if (first) {
                        const double d0 = vl_slope(vn[2], vn[3], vn[4]);
                        zsl[n] = vl_slope(vn[3], vn[4], vn[5]);
                        zf0[n] = face1(vn[3], vn[4], d0, zsl[n]);
                    }
                    const double d1 = vl_slope(vn[4], vn[5], vn[6]);
                    zf1[n] = face1(vn[4], vn[5], zsl[n], d1);
                    zsl[n] = d1;
                    sm = zf0[n];
                    sp = zf1[n];
                    cw_limit(vn[4], sm, sp);
                } else {
                    if (first) {
                        zf0[n] = cs_face(vn[1], vn[2], vn[3], vn[4]);
                        zf1[n] = cs_face(vn[2], vn[3], vn[4], vn[5]);
                        zf2[n] = cs_face(vn[3], vn[4], vn[5], vn[6]);
                    }
                    const double f3 = cs_face(vn[4], vn[5], vn[6], vn[7]);
                    cs_cell(vn, zf0[n], zf1[n], zf2[n], f3, 1.e-10, sm, sp);
                    zf0[n] = zf1[n];
                    zf1[n] = zf2[n];
                    zf2[n] = f3;
                }
                if (ppm_special<PPM>(q, dlz, dhz, A.bclo[n][2], A.bchi[n][2]))
                    ppm_edges<PPM, 2>(vn, q, dlz, dhz, A.bclo[n][2], A.bchi[n][2], sm, sp);
                if (PPM == 1) zf0[n] = zf1[n];
                trace_sig(vn[4], sm, sp, uf_q[2], uf_q[2], sgz, sgz, rel_eps, Ipz[n], Imz[n]);
            }
            if (zface) {
#pragma unroll
                for (int n = 0; n < 3; ++n) {
                    zl[n] = Ipz_prev[n];
                    zr[n] = Imz[n];
                    vf_iface_bc(q, dlz, dhz, pzl, pzh, n == 2, v[n][3], v[n][4], zl[n], zr[n]);
                }
                uz_q[0] = vf_upwind(zl[0], zr[0], trz_q, rel_eps);
                uz_q[1] = vf_upwind(zl[1], zr[1], trz_q, rel_eps);
            }
#pragma unroll
            for (int n = 0; n < 3; ++n) Ipz_prev[n] = Ipz[n];
        }
        __syncthreads();

        // ---- B2: PPM of plane q in x and y
        if (col) {
            const double sgx = dabs(uf_q[0]) * dt / hx;
            const double sgy = dabs(uf_q[1]) * dt / hy;
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                double sm, sp, Ip, Im;
                double w9[9];
#pragma unroll
                for (int o = -4; o <= 4; ++o)
                    w9[4 + o] = (o >= -G && o <= G) ? S[n][si + o] : 0.0;
                vf_cell_edges<PPM, 0>(w9, FX[n] + (a + FO) + b * FXW, 1, I, dlx, dhx,
                                      A.bclo[n][0], A.bchi[n][0], !XYF, sm, sp);
                trace_sig(w9[4], sm, sp, uf_q[0], uf_q[0], sgx, sgx, rel_eps, Ip, Im);
                P[4 * n][tid] = Ip;
                P[4 * n + 1][tid] = Im;
#pragma unroll
                for (int o = -4; o <= 4; ++o)
                    w9[4 + o] = (o >= -G && o <= G) ? S[n][si + o * SW] : 0.0;
                vf_cell_edges<PPM, 1>(w9, FY[n] + a + (b + FO) * W, W, J, dly, dhy,
                                      A.bclo[n][1], A.bchi[n][1], !XYF, sm, sp);
                trace_sig(w9[4], sm, sp, uf_q[1], uf_q[1], sgy, sgy, rel_eps, Ip, Im);
                P[4 * n + 2][tid] = Ip;
                P[4 * n + 3][tid] = Im;
            }
        }
        __syncthreads();

        // ---- C: interface and upwind states on the x- and y-faces of plane q
        //      (FX[0..1] <- uimh(0, 1), uimh(0, 2); FY[0..1] <- uimh(1, 0), uimh(1, 2))
        if (col && a >= 1) {
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                xl[n] = P[4 * n][tid - 1];
                xr[n] = P[4 * n + 1][tid];
                if (!XYF)
                    vf_iface_bc(I, dlx, dhx, pxl, pxh, n == 0, S[n][si - 1], S[n][si], xl[n],
                                xr[n]);
            }
            const double t = ULq[ui];
            FX[0][tid] = vf_upwind(xl[1], xr[1], t, rel_eps);
            FX[1][tid] = vf_upwind(xl[2], xr[2], t, rel_eps);
        }
        if (col && b >= 1) {
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                yl[n] = P[4 * n + 2][tid - W];
                yr[n] = P[4 * n + 3][tid];
                if (!XYF)
                    vf_iface_bc(J, dly, dhy, pyl, pyh, n == 1, S[n][si - SW], S[n][si], yl[n],
                                yr[n]);
            }
            const double t = VLq[vi];
            FY[0][tid] = vf_upwind(yl[0], yr[0], t, rel_eps);
            FY[1][tid] = vf_upwind(yl[2], yr[2], t, rel_eps);
        }
        __syncthreads();

        // ---- D1: per-cell transverse corrections
        //      P[0] <- (y; w) plane q, P[1] <- (x; w) plane q,
        //      P[2] <- (z; v) plane p, P[3] <- (z; u) plane p
        if (col) {
            if (a >= 1 && a <= nx) {
                const double s = ULq[ui + 1] + ULq[ui];
                t01_q = t6x * s * (FX[0][tid + 1] - FX[0][tid]);       // (x; v)
                P[1][tid] = t6x * s * (FX[1][tid + 1] - FX[1][tid]);  // (x; w)
            }
            if (b >= 1 && b <= ny) {
                const double s = VLq[vi + W] + VLq[vi];
                t10_q = t6y * s * (FY[0][tid + W] - FY[0][tid]);       // (y; u)
                P[0][tid] = t6y * s * (FY[1][tid + W] - FY[1][tid]);  // (y; w)
            }
            if (pplane) {
                const double s = trz_q + trz_p;
                P[2][tid] = t6z * s * (uz_q[1] - uz_p[1]);  // (z; v)
                P[3][tid] = t6z * s * (uz_q[0] - uz_p[0]);  // (z; u)
            }
        }
        __syncthreads();

        // ---- D2: transverse states
        //      FX[2] <- wimhxy (x-faces, plane q), FY[2] <- wimhyx (y-faces, q),
        //      P[4] <- vimhxz (x-faces, plane p), P[5] <- uimhyz (y-faces, p)
        if (col && a >= 1 && b >= 1 && b <= ny) {  // wimhxy: w on x-faces, by y
            double l = xl[2] - P[0][tid - 1];
            double r = xr[2] - P[0][tid];
            if (!XYF) vf_trans_bc(I, dlx, dhx, pxl, pxh, S[2][si - 1], S[2][si], l, r);
            FX[2][tid] = vf_upwind(l, r, ULq[ui], rel_eps);
        }
        if (col && b >= 1 && a >= 1 && a <= nx) {  // wimhyx: w on y-faces, by x
            double l = yl[2] - P[1][tid - W];
            double r = yr[2] - P[1][tid];
            if (!XYF) vf_trans_bc(J, dly, dhy, pyl, pyh, S[2][si - SW], S[2][si], l, r);
            FY[2][tid] = vf_upwind(l, r, VLq[vi], rel_eps);
        }
        if (pplane && col && a >= 1) {  // vimhxz: v on x-faces of plane p, by z
            double l = xl1_p - P[2][tid - 1];
            double r = xr1_p - P[2][tid];
            if (!XYF) vf_trans_bc(I, dlx, dhx, pxl, pxh, SC[1][tid - 1], SC[1][tid], l, r);
            P[4][tid] = vf_upwind(l, r, ULp[ui], rel_eps);
        }
        if (pplane && col && b >= 1) {  // uimhyz: u on y-faces of plane p, by z
            double l = yl0_p - P[3][tid - W];
            double r = yr0_p - P[3][tid];
            if (!XYF) vf_trans_bc(J, dly, dhy, pyl, pyh, SC[0][tid - W], SC[0][tid], l, r);
            P[5][tid] = vf_upwind(l, r, VLp[vi], rel_eps);
        }
        if (zface && col && a >= 1 && a <= nx) {  // vimhzx: v on z-face q, by x
            double l = zl[1] - t01_p;
            double r = zr[1] - t01_q;
            vf_trans_bc(q, dlz, dhz, pzl, pzh, v[1][3], v[1][4], l, r);
            p201_q = vf_upwind(l, r, trz_q, rel_eps);
        }
        if (zface && col && b >= 1 && b <= ny) {  // uimhzy: u on z-face q, by y
            double l = zl[0] - t10_p;
            double r = zr[0] - t10_q;
            vf_trans_bc(q, dlz, dhz, pzl, pzh, v[0][3], v[0][4], l, r);
            p210_q = vf_upwind(l, r, trz_q, rel_eps);
        }
        __syncthreads();

        // ---- E1: per-cell final corrections
        //      P[6], P[7] <- umac's (y), (z) of plane p; P[8], P[9] <- vmac's (x), (z)
        if (col && a >= 1 && a <= nx && b >= 1 && b <= ny) {  // wmac's, plane q
            ez0_q = t4x * (ULq[ui + 1] + ULq[ui]) * (FX[2][tid + 1] - FX[2][tid]);
            ez1_q = t4y * (VLq[vi + W] + VLq[vi]) * (FY[2][tid + W] - FY[2][tid]);
        }
        if (pplane && col) {
            const double s = trz_q + trz_p;
            if (b >= 1 && b <= ny) {  // umac's: uimhyz across y, uimhzy across z
                P[6][tid] = t4y * (VLp[vi + W] + VLp[vi]) * (P[5][tid + W] - P[5][tid]);
                P[7][tid] = t4z * s * (p210_q - p210_p);
            }
            if (a >= 1 && a <= nx) {  // vmac's: vimhxz across x, vimhzx across z
                P[8][tid] = t4x * (ULp[ui + 1] + ULp[ui]) * (P[4][tid + 1] - P[4][tid]);
                P[9][tid] = t4z * s * (p201_q - p201_p);
            }
        }
        __syncthreads();

        // ---- E2: MAC velocities
        if (pplane && col && a >= 1 && b >= 1 && b <= ny && (a <= nx || last_x)) {  // umac, p
            double vl = xl0_p - P[6][tid - 1];
            vl = vl - P[7][tid - 1];
            double vr = xr0_p - P[6][tid];
            vr = vr - P[7][tid];
            const double umacl = vl + dt2 * FL[0][tid - 1];
            const double umacr = vr + dt2 * FL[0][tid];
            const double w = sph ? B.w0m[0](I, J, q - 1) : 0.0;
            double res = vf_riemann(umacl, umacr, sph, w, rel_eps);
            if (!XYF)
                res = vf_vel_bc(res, I, dlx, dhx, pxl, pxh, SC[0][tid - 1], SC[0][tid], umacl,
                                umacr);
            oxCol[-oxks] = res;  // plane p
        }
        if (pplane && col && b >= 1 && a >= 1 && a <= nx && (b <= ny || last_y)) {  // vmac, p
            double vl = yl1_p - P[8][tid - W];
            vl = vl - P[9][tid - W];
            double vr = yr1_p - P[8][tid];
            vr = vr - P[9][tid];
            const double umacl = vl + dt2 * FL[1][tid - W];
            const double umacr = vr + dt2 * FL[1][tid];
            const double w = sph ? B.w0m[1](I, J, q - 1) : 0.0;
            double res = vf_riemann(umacl, umacr, sph, w, rel_eps);
            if (!XYF)
                res = vf_vel_bc(res, J, dly, dhy, pyl, pyh, SC[1][tid - W], SC[1][tid], umacl,
                                umacr);
            oyCol[-oyks] = res;  // plane p
        }
        if (((zface && q <= z1) || (q == z1 + 1 && last_z)) && col && a >= 1 && a <= nx &&
            b >= 1 && b <= ny) {  // wmac, z-face q
            double vl = zl[2] - ez0_p;
            vl = vl - ez1_p;
            double vr = zr[2] - ez0_q;
            vr = vr - ez1_q;
            const double umacl = vl + dt2 * foz_p;
            const double umacr = vr + dt2 * fo_q[2];
            // planar: w0_cart of the cell at the face's index (MaestroVelPred.cpp:1880)
            const double w = sph ? B.w0m[2](I, J, q) : B.w0c(I, J, q, 2);
            double res = vf_riemann(umacl, umacr, true, w, rel_eps);
            res = vf_vel_bc(res, q, dlz, dhz, pzl, pzh, v[2][3], v[2][4], umacl, umacr);
            *ozCol = res;  // z-face q
        }

        // ---- advance the column state to the next plane
        xl0_p = xl[0];
        xr0_p = xr[0];
        xl1_p = xl[1];
        xr1_p = xr[1];
        yl0_p = yl[0];
        yr0_p = yr[0];
        yl1_p = yl[1];
        yr1_p = yr[1];
        uz_p[0] = uz_q[0];
        uz_p[1] = uz_q[1];
        t01_p = t01_q;
        t10_p = t10_q;
        p201_p = p201_q;
        p210_p = p210_q;
        ez0_p = ez0_q;
        ez1_p = ez1_q;
        oxCol += oxks;
        oyCol += oyks;
        ozCol += ozks;
        slot_p = slot_q;
        slot_q = slot_q == 2 ? 0 : slot_q + 1;
    }
}

// ---------------------------------------------------------------------------
// One velocity component per block.  umac on N-faces depends on component N
// alone: its interface states iface(N, N), the transverse states of
// component N (pairs (T, 3 - N - T)), their corrections (uimh(T', N)) and
// the N-component force.  So the blocks of blockIdx.z = N compute the chain
// of component N only and write umac_N: the column state of one component
// instead of three (the staged kernels' work, split by output).
//   N = 0: pairs uimhyz (y-faces of plane q - 1, by z) and uimhzy (z-face q,
//          by y); umac on the x-faces of plane q - 1;
//   N = 1: pairs vimhxz (x-faces of plane q - 1, by z) and vimhzx (z-face q,
//          by x); vmac on the y-faces of plane q - 1;
//   N = 2: pairs wimhxy (x-faces of plane q, by y) and wimhyx (y-faces of
//          plane q, by x); wmac on z-face q.
// the block's LDS, one set for the three component bodies
template <int PPM, int TX, int TY>
struct VPShared {
    static constexpr int G = PPM == 2 ? 3 : 2;
    static constexpr int W = TX + 2;
    static constexpr int H = TY + 2;
    static constexpr int NE = W * H;
    static constexpr int SW = W + 2 * G;
    static constexpr int SH = H + 2 * G;
    static constexpr int FO = PPM == 2 ? 1 : 0;
    static constexpr int FXW = W + 1 + 2 * FO;
    static constexpr int FYH = H + 1 + 2 * FO;
    static constexpr int NF = FXW * H > W * FYH ? FXW * H : W * FYH;
    double S[SW * SH];
    double UL[3][(W + 1) * H];
    double VL[3][W * (H + 1)];
    double FX[NF];    // x-face values; then uimh(0, N); then transverse states
    double FY[NF];    // y-face values; then uimh(1, N); then transverse states
    double P[4][NE];  // Ipx, Imx, Ipy, Imy; then final corrections
    double Q[2][NE];  // transverse corrections across cells
    double SC[NE];    // plane q - 1 cell values (inflow overrides)
    double FL[NE];    // plane q - 1 force (component N)
};

template <int PPM, int TX, int TY, bool XYF, int N>
__device__ __forceinline__ void vp_comp(const VFBox& B, const VFTile& tl, const KGeom& g,
                                        const VFArgs& A, VPShared<PPM, TX, TY>& sh) {
    constexpr int NT = vf_nt(TX, TY);
    constexpr int G = PPM == 2 ? 3 : 2;
    constexpr int W = TX + 2;
    constexpr int H = TY + 2;
    constexpr int NE = W * H;
    constexpr int SW = W + 2 * G;
    constexpr int SH = H + 2 * G;
    constexpr int UW = W + 1;
    constexpr int VH = H + 1;
    constexpr int FO = PPM == 2 ? 1 : 0;
    constexpr int FXW = W + 1 + 2 * FO;
    constexpr int FYH = H + 1 + 2 * FO;
    constexpr int NF = FXW * H > W * FYH ? FXW * H : W * FYH;

    static_assert(NF == VPShared<PPM, TX, TY>::NF, "LDS layout");
    double* S = sh.S;
    auto& UL = sh.UL;
    auto& VL = sh.VL;
    double* FX = sh.FX;
    double* FY = sh.FY;
    auto& P = sh.P;
    auto& Q = sh.Q;
    double* SC = sh.SC;
    double* FL = sh.FL;

    const int bhi0 = B.hi[0], bhi1 = B.hi[1], bhi2 = B.hi[2];
    const int X0 = B.lo[0] + tl.tx * TX;
    const int Y0 = B.lo[1] + tl.ty * TY;
    const int nx = min(TX, bhi0 - X0 + 1);
    const int ny = min(TY, bhi1 - Y0 + 1);
    const int z0 = B.lo[2] + tl.tzc * A.tz;
    const int z1 = min(z0 + A.tz - 1, bhi2);
    const bool last_x = (X0 + nx - 1 == bhi0);
    const bool last_y = (Y0 + ny - 1 == bhi1);
    const bool last_z = (z1 == bhi2);
    const int dlx = B.domlo[0], dhx = B.domhi[0];
    const int dly = B.domlo[1], dhy = B.domhi[1];
    const int dlz = B.domlo[2], dhz = B.domhi[2];
    const int pxl = A.plo[0], pxh = A.phi[0], pyl = A.plo[1], pyh = A.phi[1];
    const int pzl = A.plo[2], pzh = A.phi[2];
    const bool sph = A.spherical != 0;
    const int bcx0 = A.bclo[N][0], bcx1 = A.bchi[N][0];
    const int bcy0 = A.bclo[N][1], bcy1 = A.bchi[N][1];
    const int bcz0 = A.bclo[N][2], bcz1 = A.bchi[N][2];

    const DA utv = B.ut, ufv = B.uf, fov = B.fo;
    const DA trx = B.tr[0], tryv = B.tr[1], trz = B.tr[2];
    const DA umo = B.um[N];
    const int sjs = (int)utv.js, sks = (int)utv.ks;
    const int xjs = (int)trx.js, xks = (int)trx.ks;
    const int yjs = (int)tryv.js, yks = (int)tryv.ks;

    const int tid = threadIdx.x;
    const int a = tid % W;  // local column: cell (X0 + a - 1, Y0 + b - 1)
    const int b = tid / W;
    const bool col = (tid < NE) && (a <= nx + 1) && (b <= ny + 1);
    const int I = X0 + a - 1;
    const int J = Y0 + b - 1;
    const int si = (a + G) + (b + G) * SW;
    const int ui = a + b * UW;
    const int vi = a + b * W;

    const double dt = g.dt;
    const double rel_eps = g.rel_eps;
    const double dt2 = 0.5 * dt;
    const double dt4 = 0.25 * dt;
    const double dt6 = dt / 6.0;
    const double hx = B.dx[0], hy = B.dx[1], hz = B.dx[2];
    const double t6x = dt6 / hx, t6y = dt6 / hy, t6z = dt6 / hz;
    const double t4x = dt4 / hx, t4y = dt4 / hy, t4z = dt4 / hz;

    gcd* sPl = (gcd*)utv.p + utv.off(X0 - 1 - G, Y0 - 1 - G, z0 - 1, N);
    gcd* xPl = (gcd*)trx.p + trx.off(X0 - 1, Y0 - 1, z0 - 1);
    gcd* yPl = (gcd*)tryv.p + tryv.off(X0 - 1, Y0 - 1, z0 - 1);
    gcd* sCol = (gcd*)utv.p;
    gcd* ufCol = (gcd*)ufv.p;
    gcd* foCol = (gcd*)fov.p;
    gcd* zCol = (gcd*)trz.p;
    gd* oCol = (gd*)umo.p;
    if (col) {
        sCol += utv.off(I, J, z0 - 1 + G, N);
        ufCol += ufv.off(I, J, z0 - 1);
        foCol += fov.off(I, J, z0 - 1, N);
        zCol += trz.off(I, J, z0 - 1);
        oCol += umo.off(I, J, z0 - 1);
    }
    const int ufks = (int)ufv.ks, ufns = (int)ufv.ns;
    const int foks = (int)fov.ks, zks = (int)trz.ks, oks = (int)umo.ks;

    double v[9];
#pragma unroll
    for (int jj = 0; jj < 9; ++jj) v[jj] = 0.0;
    double zf0 = 0.0, zf1 = 0.0, zf2 = 0.0, zsl = 0.0;
    double Ipz_prev = 0.0;
    double zl = 0.0, zr = 0.0;          // iface(2, N) on z-face q
    double uz_q = 0.0, uz_p = 0.0;      // uimh(2, N) on z-faces q and q - 1 (N < 2)
    double xl = 0.0, xr = 0.0, yl = 0.0, yr = 0.0;  // iface(0, N), iface(1, N), plane q
    double xl_p = 0.0, xr_p = 0.0, yl_p = 0.0, yr_p = 0.0;  // the same, plane p
    double tc_q = 0.0, tc_p = 0.0;      // own-column correction (pair on z-faces)
    double pz_q = 0.0, pz_p = 0.0;      // transverse state on z-faces q, q - 1
    double ez0_q = 0.0, ez0_p = 0.0, ez1_q = 0.0, ez1_p = 0.0;  // wmac corrections
    double trz_q = 0.0, trz_p = 0.0;
    double fo_q = 0.0, fo_p = 0.0;
    if (col)
        for (int jj = 5 - G; jj <= 4 + G; ++jj) v[jj] = utv(I, J, z0 - 6 + jj, N);
    int slot_q = 0, slot_p = 2;

    constexpr int NS = (SW * SH + NT - 1) / NT;
    constexpr int NU = (UW * H + NT - 1) / NT;
    constexpr int NV = (W * VH + NT - 1) / NT;
    bool okS[NS], okU[NU], okV[NV];
    int offS[NS], offU[NU], offV[NV];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const int idx = tid + k * NT;
        const int sx = idx % SW, sy = idx / SW;
        okS[k] = idx < SW * SH && sx <= nx + 1 + 2 * G && sy <= ny + 1 + 2 * G;
        offS[k] = sx + sy * sjs;
    }
#pragma unroll
    for (int k = 0; k < NU; ++k) {
        const int idx = tid + k * NT;
        const int f = idx % UW, r = idx / UW;
        okU[k] = idx < UW * H && f <= nx + 2 && r <= ny + 1;
        offU[k] = f + r * xjs;
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const int idx = tid + k * NT;
        const int cc = idx % W, r = idx / W;
        okV[k] = idx < W * VH && cc <= nx + 1 && r <= ny + 2;
        offV[k] = cc + r * yjs;
    }
    double pS[NS], pU[NU], pV[NV];
    double pcs = 0.0, pfo = 0.0, pz = 0.0;
#pragma unroll
    for (int k = 0; k < NS; ++k) pS[k] = okS[k] ? sPl[offS[k]] : 0.0;
#pragma unroll
    for (int k = 0; k < NU; ++k) pU[k] = okU[k] ? xPl[offU[k]] : 0.0;
#pragma unroll
    for (int k = 0; k < NV; ++k) pV[k] = okV[k] ? yPl[offV[k]] : 0.0;
    if (col) {
        pcs = *sCol;
        pfo = *foCol;
        pz = *zCol;
    }

    for (int q = z0 - 1; q <= z1 + 1; ++q) {
        const bool zface = q >= z0;
        const bool pplane = q >= z0 + 1;
        const bool first = q == z0 - 1;
        double* ULq = UL[slot_q];
        double* VLq = VL[slot_q];
        const double* ULp = UL[slot_p];
        const double* VLp = VL[slot_p];

        // ---- A: stage plane q; shift the column window; issue the loads
        //      of plane q + 1 (the tracing velocity of plane q is read here:
        //      B1 and B2 consume it)
#pragma unroll
        for (int k = 0; k < NS; ++k)
            if (okS[k]) S[tid + k * NT] = pS[k];
#pragma unroll
        for (int k = 0; k < NU; ++k)
            if (okU[k]) ULq[tid + k * NT] = pU[k];
#pragma unroll
        for (int k = 0; k < NV; ++k)
            if (okV[k]) VLq[tid + k * NT] = pV[k];
        double ufx = 0.0, ufy = 0.0, ufz = 0.0;
        if (col) {
            ufx = ufCol[0];
            ufy = ufCol[ufns];
            ufz = ufCol[2 * ufns];
#pragma unroll
            for (int jj = 0; jj < 8; ++jj) v[jj] = v[jj + 1];
            v[4 + G] = pcs;
            fo_p = fo_q;
            fo_q = pfo;
            trz_p = trz_q;
            trz_q = pz;
        }
        if (q <= z1) {
            sPl += sks;
            xPl += xks;
            yPl += yks;
#pragma unroll
            for (int k = 0; k < NS; ++k)
                if (okS[k]) pS[k] = sPl[offS[k]];
#pragma unroll
            for (int k = 0; k < NU; ++k)
                if (okU[k]) pU[k] = xPl[offU[k]];
#pragma unroll
            for (int k = 0; k < NV; ++k)
                if (okV[k]) pV[k] = yPl[offV[k]];
            if (col) {
                sCol += sks;
                foCol += foks;
                zCol += zks;
                pcs = *sCol;
                pfo = *foCol;
                pz = *zCol;
            }
        }
        if (col) ufCol += ufks;
        __syncthreads();

        // plane q - 1's cell value and force, which D2 and E2 read across
        // columns: stored after the barrier above, which no thread passes
        // before every thread is done with the previous plane's E2
        if (col) {
            SC[tid] = v[3];
            FL[tid] = fo_p;
        }

        // ---- B1: face values of plane q; PPM of cell q in z
        for (int idx = tid; idx < FXW * H; idx += NT) {
            const int f = idx % FXW - FO, r = idx / FXW;
            if (f <= nx + 2 + FO && r <= ny + 1) {
                const double* sr = S + (f + G) + (r + G) * SW;
                if (PPM == 1) {
                    const double sa2 = sr[-2], sa = sr[-1], sb = sr[0], sb2 = sr[1];
                    FX[idx] = face1(sa, sb, vl_slope(sa2, sa, sb), vl_slope(sa, sb, sb2));
                } else {
                    FX[idx] = cs_face(sr[-2], sr[-1], sr[0], sr[1]);
                }
            }
        }
        for (int idx = tid; idx < W * FYH; idx += NT) {
            const int c = idx % W, f = idx / W - FO;
            if (c <= nx + 1 && f <= ny + 2 + FO) {
                const double* sr = S + (c + G) + (f + G) * SW;
                if (PPM == 1) {
                    const double sa2 = sr[-2 * SW], sa = sr[-SW], sb = sr[0], sb2 = sr[SW];
                    FY[idx] = face1(sa, sb, vl_slope(sa2, sa, sb), vl_slope(sa, sb, sb2));
                } else {
                    FY[idx] = cs_face(sr[-2 * SW], sr[-SW], sr[0], sr[SW]);
                }
            }
        }
        if (col) {
            double sm, sp, Ipz, Imz;
            if (PPM == 1) {
                if (first) {
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
                if (first) {
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
            if (ppm_special<PPM>(q, dlz, dhz, bcz0, bcz1))
                ppm_edges<PPM, 2>(v, q, dlz, dhz, bcz0, bcz1, sm, sp);
            if (PPM == 1) zf0 = zf1;
            const double sgz = dabs(ufz) * dt / hz;
            trace_sig(v[4], sm, sp, ufz, ufz, sgz, sgz, rel_eps, Ipz, Imz);
            if (zface) {
                zl = Ipz_prev;
                zr = Imz;
                vf_iface_bc(q, dlz, dhz, pzl, pzh, N == 2, v[3], v[4], zl, zr);
                if (N != 2) uz_q = vf_upwind(zl, zr, trz_q, rel_eps);
            }
            Ipz_prev = Ipz;
        }
        __syncthreads();

        // ---- B2: PPM of plane q in x and y
        if (col) {
            double sm, sp, Ip, Im;
            double w9[9];
#pragma unroll
            for (int o = -4; o <= 4; ++o) w9[4 + o] = (o >= -G && o <= G) ? S[si + o] : 0.0;
            vf_cell_edges<PPM, 0>(w9, FX + (a + FO) + b * FXW, 1, I, dlx, dhx, bcx0, bcx1, !XYF,
                                  sm, sp);
            const double sgx = dabs(ufx) * dt / hx;
            trace_sig(w9[4], sm, sp, ufx, ufx, sgx, sgx, rel_eps, Ip, Im);
            P[0][tid] = Ip;
            P[1][tid] = Im;
#pragma unroll
            for (int o = -4; o <= 4; ++o) w9[4 + o] = (o >= -G && o <= G) ? S[si + o * SW] : 0.0;
            vf_cell_edges<PPM, 1>(w9, FY + a + (b + FO) * W, W, J, dly, dhy, bcy0, bcy1, !XYF,
                                  sm, sp);
            const double sgy = dabs(ufy) * dt / hy;
            trace_sig(w9[4], sm, sp, ufy, ufy, sgy, sgy, rel_eps, Ip, Im);
            P[2][tid] = Ip;
            P[3][tid] = Im;
        }
        __syncthreads();

        // ---- C: interface states on the x- and y-faces of plane q; the
        //      upwind states uimh(0, N) -> FX, uimh(1, N) -> FY where used
        if (col && a >= 1) {
            xl = P[0][tid - 1];
            xr = P[1][tid];
            if (!XYF) vf_iface_bc(I, dlx, dhx, pxl, pxh, N == 0, S[si - 1], S[si], xl, xr);
            if (N != 0) FX[tid] = vf_upwind(xl, xr, ULq[ui], rel_eps);
        }
        if (col && b >= 1) {
            yl = P[2][tid - W];
            yr = P[3][tid];
            if (!XYF) vf_iface_bc(J, dly, dhy, pyl, pyh, N == 1, S[si - SW], S[si], yl, yr);
            if (N != 1) FY[tid] = vf_upwind(yl, yr, VLq[vi], rel_eps);
        }
        __syncthreads();

        // ---- D1: per-cell corrections of component N
        if (col) {
            if (N == 2) {
                if (b >= 1 && b <= ny)  // (y; w): wimhxy on the x-faces beside it
                    Q[0][tid] = t6y * (VLq[vi + W] + VLq[vi]) * (FY[tid + W] - FY[tid]);
                if (a >= 1 && a <= nx)  // (x; w): wimhyx on the y-faces beside it
                    Q[1][tid] = t6x * (ULq[ui + 1] + ULq[ui]) * (FX[tid + 1] - FX[tid]);
            } else if (N == 1) {
                if (a >= 1 && a <= nx)  // (x; v): vimhzx on its z-faces
                    tc_q = t6x * (ULq[ui + 1] + ULq[ui]) * (FX[tid + 1] - FX[tid]);
                if (pplane)  // (z; v) of plane p: vimhxz on the x-faces beside it
                    Q[0][tid] = t6z * (trz_q + trz_p) * (uz_q - uz_p);
            } else {
                if (b >= 1 && b <= ny)  // (y; u): uimhzy on its z-faces
                    tc_q = t6y * (VLq[vi + W] + VLq[vi]) * (FY[tid + W] - FY[tid]);
                if (pplane)  // (z; u) of plane p: uimhyz on the y-faces beside it
                    Q[0][tid] = t6z * (trz_q + trz_p) * (uz_q - uz_p);
            }
        }
        __syncthreads();

        // ---- D2: transverse states of component N
        if (N == 2) {
            if (col && a >= 1 && b >= 1 && b <= ny) {  // wimhxy, x-faces of plane q
                double l = xl - Q[0][tid - 1];
                double r = xr - Q[0][tid];
                if (!XYF) vf_trans_bc(I, dlx, dhx, pxl, pxh, S[si - 1], S[si], l, r);
                FX[tid] = vf_upwind(l, r, ULq[ui], rel_eps);
            }
            if (col && b >= 1 && a >= 1 && a <= nx) {  // wimhyx, y-faces of plane q
                double l = yl - Q[1][tid - W];
                double r = yr - Q[1][tid];
                if (!XYF) vf_trans_bc(J, dly, dhy, pyl, pyh, S[si - SW], S[si], l, r);
                FY[tid] = vf_upwind(l, r, VLq[vi], rel_eps);
            }
        } else if (N == 1) {
            if (pplane && col && a >= 1) {  // vimhxz, x-faces of plane p
                double l = xl_p - Q[0][tid - 1];
                double r = xr_p - Q[0][tid];
                if (!XYF) vf_trans_bc(I, dlx, dhx, pxl, pxh, SC[tid - 1], SC[tid], l, r);
                FX[tid] = vf_upwind(l, r, ULp[ui], rel_eps);
            }
            if (zface && col && a >= 1 && a <= nx) {  // vimhzx, z-face q
                double l = zl - tc_p;
                double r = zr - tc_q;
                vf_trans_bc(q, dlz, dhz, pzl, pzh, v[3], v[4], l, r);
                pz_q = vf_upwind(l, r, trz_q, rel_eps);
            }
        } else {
            if (pplane && col && b >= 1) {  // uimhyz, y-faces of plane p
                double l = yl_p - Q[0][tid - W];
                double r = yr_p - Q[0][tid];
                if (!XYF) vf_trans_bc(J, dly, dhy, pyl, pyh, SC[tid - W], SC[tid], l, r);
                FY[tid] = vf_upwind(l, r, VLp[vi], rel_eps);
            }
            if (zface && col && b >= 1 && b <= ny) {  // uimhzy, z-face q
                double l = zl - tc_p;
                double r = zr - tc_q;
                vf_trans_bc(q, dlz, dhz, pzl, pzh, v[3], v[4], l, r);
                pz_q = vf_upwind(l, r, trz_q, rel_eps);
            }
        }
        __syncthreads();

        // ---- E1: per-cell final corrections
        if (N == 2) {
            if (col && a >= 1 && a <= nx && b >= 1 && b <= ny) {
                ez0_q = t4x * (ULq[ui + 1] + ULq[ui]) * (FX[tid + 1] - FX[tid]);
                ez1_q = t4y * (VLq[vi + W] + VLq[vi]) * (FY[tid + W] - FY[tid]);
            }
        } else if (N == 1) {
            if (pplane && col && a >= 1 && a <= nx) {  // vmac's: vimhxz across x, vimhzx across z
                P[0][tid] = t4x * (ULp[ui + 1] + ULp[ui]) * (FX[tid + 1] - FX[tid]);
                P[1][tid] = t4z * (trz_q + trz_p) * (pz_q - pz_p);
            }
        } else {
            if (pplane && col && b >= 1 && b <= ny) {  // umac's: uimhyz across y, uimhzy across z
                P[0][tid] = t4y * (VLp[vi + W] + VLp[vi]) * (FY[tid + W] - FY[tid]);
                P[1][tid] = t4z * (trz_q + trz_p) * (pz_q - pz_p);
            }
        }
        __syncthreads();

        // ---- E2: the MAC velocity of component N
        if (N == 0) {
            if (pplane && col && a >= 1 && b >= 1 && b <= ny && (a <= nx || last_x)) {
                double vl = xl_p - P[0][tid - 1];
                vl = vl - P[1][tid - 1];
                double vr = xr_p - P[0][tid];
                vr = vr - P[1][tid];
                const double umacl = vl + dt2 * FL[tid - 1];
                const double umacr = vr + dt2 * FL[tid];
                const double w = sph ? B.w0m[0](I, J, q - 1) : 0.0;
                double res = vf_riemann(umacl, umacr, sph, w, rel_eps);
                if (!XYF)
                    res = vf_vel_bc(res, I, dlx, dhx, pxl, pxh, SC[tid - 1], SC[tid], umacl,
                                    umacr);
                oCol[-oks] = res;  // plane p
            }
        } else if (N == 1) {
            if (pplane && col && b >= 1 && a >= 1 && a <= nx && (b <= ny || last_y)) {
                double vl = yl_p - P[0][tid - W];
                vl = vl - P[1][tid - W];
                double vr = yr_p - P[0][tid];
                vr = vr - P[1][tid];
                const double umacl = vl + dt2 * FL[tid - W];
                const double umacr = vr + dt2 * FL[tid];
                const double w = sph ? B.w0m[1](I, J, q - 1) : 0.0;
                double res = vf_riemann(umacl, umacr, sph, w, rel_eps);
                if (!XYF)
                    res = vf_vel_bc(res, J, dly, dhy, pyl, pyh, SC[tid - W], SC[tid], umacl,
                                    umacr);
                oCol[-oks] = res;  // plane p
            }
        } else {
            if (((zface && q <= z1) || (q == z1 + 1 && last_z)) && col && a >= 1 && a <= nx &&
                b >= 1 && b <= ny) {
                double vl = zl - ez0_p;
                vl = vl - ez1_p;
                double vr = zr - ez0_q;
                vr = vr - ez1_q;
                const double umacl = vl + dt2 * fo_p;
                const double umacr = vr + dt2 * fo_q;
                // planar: w0_cart of the cell at the face's index (MaestroVelPred.cpp:1880)
                const double w = sph ? B.w0m[2](I, J, q) : B.w0c(I, J, q, 2);
                double res = vf_riemann(umacl, umacr, true, w, rel_eps);
                res = vf_vel_bc(res, q, dlz, dhz, pzl, pzh, v[3], v[4], umacl, umacr);
                *oCol = res;  // z-face q
            }
        }

        // ---- advance the column state to the next plane
        xl_p = xl;
        xr_p = xr;
        yl_p = yl;
        yr_p = yr;
        uz_p = uz_q;
        tc_p = tc_q;
        pz_p = pz_q;
        ez0_p = ez0_q;
        ez1_p = ez1_q;
        oCol += oks;
        slot_p = slot_q;
        slot_q = slot_q == 2 ? 0 : slot_q + 1;
    }
}

// xmode 1, 2: the XCD-aware orders of xcd_tile over a 1-D grid; 0: tile
// blockIdx.x and component blockIdx.z
template <int PPM, int TX, int TY, bool XYF, int WPE>
__global__ void __launch_bounds__(vf_nt(TX, TY)) __attribute__((amdgpu_waves_per_eu(WPE)))
    k_vp_fused_comp(const VFBox* __restrict__ boxes, const VFTile* __restrict__ tiles, KGeom g,
                    VFArgs A, int ntile, int xmode) {
    __shared__ VPShared<PPM, TX, TY> sh;
    int t = blockIdx.x, n = blockIdx.z;
    if (xmode != 0 && !xcd_tile(ntile, 3, xmode, t, n)) return;
    const VFTile tl = tiles[t];
    const VFBox& B = boxes[tl.box];
    if (n == 0)
        vp_comp<PPM, TX, TY, XYF, 0>(B, tl, g, A, sh);
    else if (n == 1)
        vp_comp<PPM, TX, TY, XYF, 1>(B, tl, g, A, sh);
    else
        vp_comp<PPM, TX, TY, XYF, 2>(B, tl, g, A, sh);
}

}  // namespace

namespace mxh {

bool velpred_fused_ok(const mx_adv_ctx* c) {
    return c->p.dim == 3 && (c->p.ppm_type == 1 || c->p.ppm_type == 2) &&
           c->p.ppm_trace_forces == 0 && c->vp_fused;
}

namespace {

bool phys_acts(int bc) { return bc != MX_PHYS_INTERIOR; }

bool bc_type_acts(int bc) { return bc == MX_BC_EXT_DIR || bc == MX_BC_HOEXTRAP; }

// x/y-free: in x and in y each domain side either carries nothing that acts
// (no physical override, no PPM boundary stencil) or lies beyond the tile's
// reach (cells [X0 - 1, X0 + nx], faces [X0, X0 + nx]; stencils act within
// three cells of a side)
bool vf_tile_free(int lo0, int n, int dlo, int dhi, bool act_lo, bool act_hi) {
    const bool free_lo = !act_lo || lo0 - 1 > dlo + 2;
    const bool free_hi = !act_hi || lo0 + n < dhi - 2;
    return free_lo && free_hi;
}

// COMP: one velocity component per block (blockIdx.z), else all three
template <int TX, int TY, int WPE, bool COMP, bool XYF>
void vf_launch_one(mx_adv_ctx* c, const VFBox* db, const VFTile* dt, size_t n, const KGeom& kg,
                   const VFArgs& A, int xm, int& rc) {
    if (n == 0 || rc) return;
    auto launch = [&]() -> int {
        if constexpr (COMP) {
            const dim3 grid = xm ? dim3(xcd_tile_blocks((long long)n, 3)) : dim3((unsigned)n, 1, 3);
            const int nt = (int)n;
            if (c->p.ppm_type == 1)
                MX_LAUNCH(c, "k_vp_fused", (k_vp_fused_comp<1, TX, TY, XYF, WPE>), grid,
                          dim3(vf_nt(TX, TY)), db, dt, kg, A, nt, xm);
            else
                MX_LAUNCH(c, "k_vp_fused", (k_vp_fused_comp<2, TX, TY, XYF, WPE>), grid,
                          dim3(vf_nt(TX, TY)), db, dt, kg, A, nt, xm);
        } else {
            if (c->p.ppm_type == 1)
                MX_LAUNCH(c, "k_vp_fused", (k_vp_fused<1, TX, TY, XYF, WPE>), dim3((unsigned)n),
                          dim3(vf_nt(TX, TY)), db, dt, kg, A);
            else
                MX_LAUNCH(c, "k_vp_fused", (k_vp_fused<2, TX, TY, XYF, WPE>), dim3((unsigned)n),
                          dim3(vf_nt(TX, TY)), db, dt, kg, A);
        }
        return MX_OK;
    };
    rc = launch();
}

// WPE: the register budget (waves per SIMD) of the x/y-free tiles, WPEB that
// of the tiles at an acting x/y side, whose variant carries the side rules
template <int TX, int TY, int WPE, bool COMP, int WPEB = WPE>
int vf_launch(mx_adv_ctx* c, const std::vector<VpLevel>& levs, int tz) {
    const long long lim = 0x7fffffffLL;
    const bool sph = c->p.spherical != 0;
    std::vector<VFBox> boxes;
    std::vector<VFTile> tiles, btiles;
    bool act[2][2] = {{false, false}, {false, false}};
    for (int d = 0; d < 2; ++d) {
        act[d][0] = phys_acts(c->p.phys_bc_lo[d]);
        act[d][1] = phys_acts(c->p.phys_bc_hi[d]);
        for (int n = 0; n < 3; ++n) {
            act[d][0] = act[d][0] || bc_type_acts(c->bcs_u[n].lo[d]);
            act[d][1] = act[d][1] || bc_type_acts(c->bcs_u[n].hi[d]);
        }
    }
    for (const VpLevel& L : levs) {
        const auto& grids = c->grids[L.lev];
        for (size_t b = 0; b < grids.size(); ++b) {
            VFBox B;
            HBox g = hbox(grids[b]);
            for (int q = 0; q < 3; ++q) {
                B.lo[q] = g.lo[q];
                B.hi[q] = g.hi[q];
                B.dx[q] = c->p.dx[L.lev][q];
                B.domlo[q] = c->p.domain_lo[L.lev][q];
                B.domhi[q] = c->p.domain_hi[L.lev][q];
            }
            B.ut = to_da(L.utilde->fab[b]);
            B.uf = to_da(L.ufull->fab[b]);
            B.fo = to_da(L.force->fab[b]);
            B.w0c = (!sph && L.w0_cart) ? to_da(L.w0_cart->fab[b]) : null_da();
            for (int d = 0; d < 3; ++d) {
                B.tr[d] = to_da(L.utrans->dir[d].fab[b]);
                B.um[d] = to_da(L.umac->dir[d].fab[b]);
                B.w0m[d] = sph ? to_da(L.w0mac->dir[d].fab[b]) : null_da();
            }
            // the kernel keeps plane, row and component strides in 32 bits
            const DA* views[] = {&B.ut, &B.uf, &B.fo, &B.tr[0], &B.tr[1], &B.tr[2],
                                 &B.um[0], &B.um[1], &B.um[2]};
            for (const DA* d : views)
                if (d->js > lim || d->ks > lim || d->ns > lim) return MX_ERR_UNSUPPORTED;
            const int ntx = (g.hi[0] - g.lo[0] + TX) / TX;
            const int nty = (g.hi[1] - g.lo[1] + TY) / TY;
            const int nzc = (g.hi[2] - g.lo[2] + tz) / tz;
            const int id = (int)boxes.size();
            for (int z = 0; z < nzc; ++z)
                for (int y = 0; y < nty; ++y)
                    for (int x = 0; x < ntx; ++x) {
                        const int x0 = g.lo[0] + x * TX, y0 = g.lo[1] + y * TY;
                        const int nx = std::min(TX, g.hi[0] - x0 + 1);
                        const int ny = std::min(TY, g.hi[1] - y0 + 1);
                        const bool f =
                            vf_tile_free(x0, nx, B.domlo[0], B.domhi[0], act[0][0], act[0][1]) &&
                            vf_tile_free(y0, ny, B.domlo[1], B.domhi[1], act[1][0], act[1][1]);
                        (f ? tiles : btiles).push_back({id, x, y, z});
                    }
            boxes.push_back(B);
        }
    }
    const size_t nfree = tiles.size();
    tiles.insert(tiles.end(), btiles.begin(), btiles.end());
    if (tiles.empty()) return MX_OK;
    if (tiles.size() > 0x7fffffffULL) return MX_ERR_UNSUPPORTED;
    VFArgs A;
    for (int n = 0; n < 3; ++n)
        for (int d = 0; d < 3; ++d) {
            A.bclo[n][d] = c->bcs_u[n].lo[d];
            A.bchi[n][d] = c->bcs_u[n].hi[d];
        }
    for (int d = 0; d < 3; ++d) {
        A.plo[d] = c->p.phys_bc_lo[d];
        A.phi[d] = c->p.phys_bc_hi[d];
    }
    A.spherical = sph ? 1 : 0;
    A.tz = tz;
    Table tab(c);
    const size_t ob = tab.add(boxes);
    const size_t ot = tab.add(tiles);
    int rc = tab.upload();
    if (rc) return rc;
    const VFBox* db = tab.dev<VFBox>(ob);
    const VFTile* dtl = tab.dev<VFTile>(ot);
    const KGeom kg = kgeom(c, levs[0].lev);
    // the block order: -1 (the default) takes contiguous tile ranges per XCD
    // (2) for one level of equal boxes, whose tiles cost alike, and the
    // round robin (1) otherwise, where contiguous ranges of unequal tiles
    // unbalance the XCDs (256^3 in 8 boxes: 1.59 against 1.61 ms; the 384^3
    // star in 216 uneven boxes: 9.80 against 9.56 ms)
    int xm = c->vp_fused_xcd;
    if (xm < 0) {
        bool equal = levs.size() == 1;
        for (size_t b = 1; equal && b < boxes.size(); ++b)
            for (int q = 0; q < 3; ++q)
                equal = equal && boxes[b].hi[q] - boxes[b].lo[q] == boxes[0].hi[q] - boxes[0].lo[q];
        xm = equal ? 2 : 1;
    }
    vf_launch_one<TX, TY, WPE, COMP, true>(c, db, dtl, nfree, kg, A, xm, rc);
    vf_launch_one<TX, TY, WPEB, COMP, false>(c, db, dtl + nfree, tiles.size() - nfree, kg, A, xm,
                                             rc);
    return rc;
}

}  // namespace

int velpred_fused(mx_adv_ctx* c, const std::vector<VpLevel>& levs) {
    if (levs.empty()) return MX_OK;
    int rc = check_phys(c);
    if (rc) return rc;
    const int tz = c->vp_fused_tz > 0 ? c->vp_fused_tz : 32;
    // all three components per block: 16 x 11 columns at 2 waves per SIMD
    // (256 VGPRs); one component per block: the tile shape (cell columns)
    // and register budget (waves per SIMD) of `velpred_fused_tile`
    if (!c->vp_fused_comp) return vf_launch<16, 11, 2, false>(c, levs, tz);
    // the default, 2 (16 x 11, x/y-free tiles at 4 waves per SIMD: 128
    // VGPRs, 8 spilled for ppm_type 1, 26 for ppm_type 2), takes 1.57
    // against 1.80 ms of tile 0 per 256^3 ppm_type 1 pass; 2, 4 and 5 keep
    // the tiles at an acting x/y side at 3 waves, where 4 spill up to 107
    switch (c->vp_fused_tile) {
        case 1: return vf_launch<16, 8, 3, true>(c, levs, tz);
        case 2: return vf_launch<16, 11, 4, true, 3>(c, levs, tz);
        case 3: return vf_launch<32, 8, 3, true>(c, levs, tz);
        case 4: return vf_launch<16, 8, 4, true, 3>(c, levs, tz);
        case 5: return vf_launch<16, 12, 4, true, 3>(c, levs, tz);
        default: return vf_launch<16, 11, 3, true>(c, levs, tz);
    }
}

}  // namespace mxh
