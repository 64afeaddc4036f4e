// mx_edge_fused.hip -- fused 3-D MakeEdgeScal: PPM, predictor, transverse
// and edge states of one scalar in a single z-marching kernel.
//
// A block owns a 32 x 8 tile of cell columns of one box, grown by one cell in
// x and y (34 x 10 columns, one thread each), and marches up a chunk [z0, z1]
// of the box's planes, one cell plane q per step (plus q = z0 - 1 and
// q = z1 + 1, whose states the chunk's first and last faces need).
//
// The reference formulas (MaestroMakeEdgeScalars.cpp: Predictor :662-863,
// Transverse :865-1359, Edges :1361-1667; MaestroPPM.cpp) are regrouped, not
// changed:
//   * a PPM face value (type 1: the limited 4th-order interpolant; type 2: the
//     Colella-Sekora face value) is computed once per face and read by both
//     cells beside it; sigma = |u| dt / dx likewise, once per face;
//   * each transverse correction of a corner-coupled state,
//     dt/6h (u_T(+) + u_T(-)) (simh_T(+) - simh_T(-)), depends only on the cell
//     it is taken across, so it is computed once per cell and direction T and
//     used by every face state that crosses that cell (the left state of the
//     face above it and the right state of the face below); the edge-state
//     corrections likewise;
//   * every such term is evaluated by the reference's expression in the
//     reference's operand order, and subtracted in the reference's order, so
//     all results are bit-identical to the per-point form.
// PPM boundary stencils (ext_dir / hoextrap within two or three cells of the
// domain edge) fall back to the full per-cell routine of mx_device.H.
//
// Per step q a block runs seven phases separated by barriers:
//   A  stage plane q of s, the x- and y-MAC velocities and their sigma in LDS;
//      shift the column's z window (s, w, force) in registers;
//   B1 PPM face values of plane q (x, y) into LDS; PPM of cell q in z
//      (register window) and the upwind state simhz on z-face q;
//   B2 PPM of plane q in x and y (cell limiter + trace);
//   C  predictor states and upwind states simhx, simhy of plane q;
//   D1 per-cell transverse terms: x, y of plane q, z of plane q-1;
//   D2 corner-coupled states simhxy, simhyx (plane q), simhxz, simhyz (plane
//      q-1), simhzx, simhzy (z-face q, in registers);
//   E1 per-cell edge terms of plane q-1 (LDS) and of plane q (registers);
//   E2 edge states of plane q-1 (x and y faces) and of z-face q.
#include <algorithm>

#include "mx_internal.H"

using namespace mxd;
using namespace mxh;

namespace {

typedef __attribute__((address_space(1))) const double gcd;
typedef __attribute__((address_space(1))) double gd;

// threads of a block: the tile's columns grown by one cell on each side,
// rounded up to whole waves
__host__ __device__ constexpr int fused_nt(int tx, int ty) {
    return ((tx + 2) * (ty + 2) + 63) / 64 * 64;
}

struct FBox {
    int lo[3], hi[3];
    DA s, fo;
    DA um[3], se[3];
    // geometry of the box's level
    double dx[3];
    int domlo[3], domhi[3];
    // the flux epilogue (FComp::flux): the fluxes and the level's base
    // state at cell heights (bc_a, bc_b: old, new) and at edges (be_a, be_b)
    DA sf[3];
    const double *bc_a, *bc_b, *be_a, *be_b;
};

// one block's work: a tile of one box, chunk tzc of its planes
struct FTile {
    int box, tx, ty, tzc;
};

struct FComp {
    int bclo[3], bchi[3];
    int comp;  // state / edge-state component
    // flux epilogue of this component (MakeRhoHFlux's planar branches,
    // MaestroMakeFlux.cpp:557-620): 0 none; 1 umac (rhoh0_edge + sedge)
    // (enthalpy_pred_type 1, 3); 2 umac sedge (enthalpy_pred_type 0)
    int flux;
};

struct FArgs {
    int start_scomp;
    int num_comp;
    int is_vel;
    int tz;  // planes per chunk
};

// the flux of edge state `se` on d-face (i, j, k) (MakeRhoHFlux, planar,
// MaestroMakeFlux.cpp:557-620: umac (rhoh0_edge + sedge), the base state at
// the face's height, averaged over old and new, from the edge-centred state on
// vertical faces)
__device__ __forceinline__ void flux_store(const FBox& B, int d, int i, int j, int k, int comp,
                                           int kind, double um, double se, bool vertical) {
    double f;
    if (kind == 1) {
        const double b0 = vertical ? 0.5 * (B.be_a[k] + B.be_b[k]) : 0.5 * (B.bc_a[k] + B.bc_b[k]);
        f = um * (b0 + se);
    } else {
        f = um * se;
    }
    B.sf[d](i, j, k, comp) = f;
}

// predictor-style boundary overrides on a (left, right) pair at face `pos`
__device__ __forceinline__ void bc_pair(int pos, int dlo, int dhi, int bclo, int bchi,
                                        bool vel_normal, double sL, double sR, double& l,
                                        double& r) {
    if (pos == dlo) {
        if (bclo == MX_BC_EXT_DIR) {
            l = sL;
            r = l;
        } else if (bclo == MX_BC_FOEXTRAP || bclo == MX_BC_HOEXTRAP) {
            if (vel_normal) r = dmin(r, 0.0);
            l = r;
        } else if (bclo == MX_BC_REFLECT_EVEN) {
            l = r;
        } else if (bclo == MX_BC_REFLECT_ODD) {
            l = 0.0;
            r = 0.0;
        }
    } else if (pos == dhi + 1) {
        if (bchi == MX_BC_EXT_DIR) {
            l = sR;
            r = l;
        } else if (bchi == MX_BC_FOEXTRAP || bchi == MX_BC_HOEXTRAP) {
            if (vel_normal) l = dmax(l, 0.0);
            r = l;
        } else if (bchi == MX_BC_REFLECT_EVEN) {
            r = l;
        } else if (bchi == MX_BC_REFLECT_ODD) {
            l = 0.0;
            r = 0.0;
        }
    }
}

__device__ __forceinline__ double upwind(double l, double r, double u, double thr) {
    double res = (u > 0.0) ? l : r;
    res = (dabs(u) > thr) ? res : 0.5 * (l + r);
    return res;
}

// final edge-state override (MaestroMakeEdgeScalars.cpp:1450-1480)
__device__ __forceinline__ double bc_edge(double res, int pos, int dlo, int dhi, int bclo,
                                          int bchi, bool vel_normal, double sL, double sR,
                                          double sedgel, double sedger) {
    if (pos == dlo) {
        if (bclo == MX_BC_EXT_DIR) {
            res = sL;
        } else if (bclo == MX_BC_FOEXTRAP || bclo == MX_BC_HOEXTRAP) {
            res = vel_normal ? dmin(sedger, 0.0) : sedger;
        } else if (bclo == MX_BC_REFLECT_EVEN) {
            res = sedger;
        } else if (bclo == MX_BC_REFLECT_ODD) {
            res = 0.0;
        }
    } else if (pos == dhi + 1) {
        if (bchi == MX_BC_EXT_DIR) {
            res = sR;
        } else if (bchi == MX_BC_FOEXTRAP || bchi == MX_BC_HOEXTRAP) {
            res = vel_normal ? dmax(sedgel, 0.0) : sedgel;
        } else if (bchi == MX_BC_REFLECT_EVEN) {
            res = sedgel;
        } else if (bchi == MX_BC_REFLECT_ODD) {
            res = 0.0;
        }
    }
    return res;
}

// transverse / edge corrections of one side: x - term (advective) or
// ((x - a) - b) + c (conservative)
template <bool CONS>
__device__ __forceinline__ double apply3(double x, double a, double b, double c) {
    if (CONS) return ((x - a) - b) + c;
    return x - a;
}

// XYF: no x- or y-boundary condition acts on the tile (the host puts a tile
// in this class when its stencils keep clear of every physical x/y domain
// side, or when those sides carry no boundary override), so the per-lane
// boundary branches of those directions are left out.
// WPE: the register budget, as waves per SIMD
template <int PPM, bool CONS, int TX, int TY, bool XYF, int WPE>
__global__ void __launch_bounds__(fused_nt(TX, TY)) __attribute__((amdgpu_waves_per_eu(WPE)))
    k_es_fused(const FBox* __restrict__ boxes, const FTile* __restrict__ tiles,
               const FComp* __restrict__ comps, KGeom g, FArgs A, int ntile, int xmode) {
    constexpr int NT = fused_nt(TX, TY);
    constexpr int G = PPM == 2 ? 3 : 2;
    constexpr int W = TX + 2;
    constexpr int H = TY + 2;
    constexpr int NE = W * H;
    constexpr int SW = W + 2 * G;
    constexpr int SH = H + 2 * G;
    constexpr int UW = W + 1;  // x-faces per row of the x-velocity plane
    constexpr int VH = H + 1;  // y-faces per column of the y-velocity plane
    constexpr int FO = PPM == 2 ? 1 : 0;
    constexpr int FXW = W + 1 + 2 * FO;  // x-face values per row
    constexpr int FYH = H + 1 + 2 * FO;  // y-face values per column
    constexpr int NP2 = CONS ? 8 * NE : 4 * NE;
    static_assert(NE <= NT, "tile too large for the block");
    static_assert(FXW * H + W * FYH <= 4 * NE, "face values do not fit the pool");

    __shared__ double S[SW * SH];
    __shared__ double UL[3][UW * H];
    __shared__ double VL[3][W * VH];
    __shared__ double SGX[UW * H];
    __shared__ double SGY[W * VH];
    __shared__ double SC[NE];
    __shared__ double FL[NE];
    __shared__ double P1[4 * NE];
    __shared__ double P2[NP2];

    // xmode 1, 2: the XCD-aware orders of xcd_tile over a 1-D grid; 0: tile
    // blockIdx.x and component blockIdx.z
    int tix = blockIdx.x, vc = blockIdx.z;
    if (xmode != 0 && !xcd_tile(ntile, A.num_comp, xmode, tix, vc)) return;
    const FTile tl = tiles[tix];
    const int bid = tl.box;
    const int blo0 = boxes[bid].lo[0], blo1 = boxes[bid].lo[1], blo2 = boxes[bid].lo[2];
    const int bhi0 = boxes[bid].hi[0], bhi1 = boxes[bid].hi[1], bhi2 = boxes[bid].hi[2];
    const int tx = tl.tx;
    const int ty = tl.ty;
    const int tzc = tl.tzc;
    const int X0 = blo0 + tx * TX;
    const int Y0 = blo1 + ty * TY;
    const int nx = min(TX, bhi0 - X0 + 1);
    const int ny = min(TY, bhi1 - Y0 + 1);
    const int z0 = blo2 + tzc * A.tz;
    const int z1 = min(z0 + A.tz - 1, bhi2);
    const bool last_x = (X0 + nx - 1 == bhi0);
    const bool last_y = (Y0 + ny - 1 == bhi1);
    const bool last_z = (z1 == bhi2);

    const int comp = comps[vc].comp;
    const int flux = comps[vc].flux;
    const int bcx0 = comps[vc].bclo[0], bcx1 = comps[vc].bchi[0];
    const int bcy0 = comps[vc].bclo[1], bcy1 = comps[vc].bchi[1];
    const int bcz0 = comps[vc].bclo[2], bcz1 = comps[vc].bchi[2];
    const int dlx = boxes[bid].domlo[0], dhx = boxes[bid].domhi[0];
    const int dly = boxes[bid].domlo[1], dhy = boxes[bid].domhi[1];
    const int dlz = boxes[bid].domlo[2], dhz = boxes[bid].domhi[2];

    // views (the strides fit 32 bits; checked on the host)
    const DA sv = boxes[bid].s;
    const DA fv = boxes[bid].fo;
    const DA uv = boxes[bid].um[0];
    const DA vv = boxes[bid].um[1];
    const DA wv = boxes[bid].um[2];
    const DA e0 = boxes[bid].se[0];
    const DA e1 = boxes[bid].se[1];
    const DA e2 = boxes[bid].se[2];
    const int sjs = (int)sv.js, sks = (int)sv.ks;
    const int ujs = (int)uv.js, uks = (int)uv.ks;
    const int vjs = (int)vv.js, vks = (int)vv.ks;

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
    const double dt3 = dt / 3.0;
    const double dt4 = 0.25 * dt;
    const double dt6 = dt / 6.0;
    const double hx = boxes[bid].dx[0], hy = boxes[bid].dx[1], hz = boxes[bid].dx[2];
    // per-direction factors of the transverse (t) and edge (e) terms
    const double tfx = CONS ? dt3 / hx : dt6 / hx;
    const double tfy = CONS ? dt3 / hy : dt6 / hy;
    const double tfz = CONS ? dt3 / hz : dt6 / hz;
    const double efx = CONS ? dt2 / hx : dt4 / hx;
    const double efy = CONS ? dt2 / hy : dt4 / hy;
    const double efz = CONS ? dt2 / hz : dt4 / hz;
    const double dt2hx = dt2 / hx, dt2hy = dt2 / hy, dt2hz = dt2 / hz;
    const bool vn0 = A.is_vel && comp == 0;
    const bool vn1 = A.is_vel && comp == 1;
    const bool vn2 = A.is_vel && comp == 2;

    // plane pointers (block-uniform) at plane z0 - 1
    gcd* sPl = (gcd*)sv.p + sv.off(X0 - 1 - G, Y0 - 1 - G, z0 - 1, comp);
    gcd* uPl = (gcd*)uv.p + uv.off(X0 - 1, Y0 - 1, z0 - 1);
    gcd* vPl = (gcd*)vv.p + vv.off(X0 - 1, Y0 - 1, z0 - 1);
    // column pointers
    gcd* sCol = (gcd*)sv.p;
    gcd* fCol = (gcd*)fv.p;
    gcd* wCol = (gcd*)wv.p;
    gd* e0Col = (gd*)e0.p;
    gd* e1Col = (gd*)e1.p;
    gd* e2Col = (gd*)e2.p;
    if (col) {
        sCol += sv.off(I, J, z0 - 1 + G, comp);  // next z-window load: s(q + G)
        fCol += fv.off(I, J, z0 - 1, comp);
        wCol += wv.off(I, J, z0 - 1);
        e0Col += e0.off(I, J, z0 - 1, comp);
        e1Col += e1.off(I, J, z0 - 1, comp);
        e2Col += e2.off(I, J, z0 - 1, comp);
    }
    const int fks = (int)fv.ks, wks = (int)wv.ks;
    const int e0ks = (int)e0.ks, e1ks = (int)e1.ks, e2ks = (int)e2.ks;

    // column state
    double v[9];
    for (int jj = 0; jj < 9; ++jj) v[jj] = 0.0;
    double w_q = 0.0, w_p = 0.0, w_hi = 0.0, sig_lo = 0.0;
    double fo_q = 0.0, fo_p = 0.0;
    double zf0 = 0.0, zf1 = 0.0, zf2 = 0.0, zsl = 0.0;  // z face values / slope queue
    double Ipz_prev = 0.0, slz = 0.0, srz = 0.0, hz_q = 0.0, hz_p = 0.0;
    double slx = 0.0, srx = 0.0, sly = 0.0, sry = 0.0;
    double slx_p = 0.0, srx_p = 0.0, sly_p = 0.0, sry_p = 0.0;
    double tx_q = 0.0, tx_p = 0.0, ty_q = 0.0, ty_p = 0.0;     // own transverse terms
    double cx_q = 0.0, cx_p = 0.0, cy_q = 0.0, cy_p = 0.0;     // conservative parts
    double bq = 0.0, bp = 0.0;
    double tzx_q = 0.0, tzx_p = 0.0, tzy_q = 0.0, tzy_p = 0.0;  // simhzx, simhzy
    double exy_q = 0.0, exy_p = 0.0, eyx_q = 0.0, eyx_p = 0.0;  // own edge terms
    if (col) {
        for (int jj = 5 - G; jj <= 4 + G; ++jj) v[jj] = sv(I, J, z0 - 6 + jj, comp);
        w_q = wv(I, J, z0 - 1);
        sig_lo = dabs(w_q) * dt / hz;
    }
    int slot_q = 0, slot_p = 2;

    // Plane q + 1's loads are issued during step q and land in registers
    // (the barriers wait for LDS traffic only), so the global-memory latency
    // overlaps the step's work.  Each thread owns fixed slots of the staged
    // planes: s (NS per thread), the x- and y-velocities (NU, NV).
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
        offU[k] = f + r * ujs;
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const int idx = tid + k * NT;
        const int cc = idx % W, r = idx / W;
        okV[k] = idx < W * VH && cc <= nx + 1 && r <= ny + 2;
        offV[k] = cc + r * vjs;
    }
    double pS[NS], pU[NU], pV[NV];
    double pcs = 0.0, pw = 0.0, pf = 0.0;
    // plane z0 - 1
#pragma unroll
    for (int k = 0; k < NS; ++k) pS[k] = okS[k] ? sPl[offS[k]] : 0.0;
#pragma unroll
    for (int k = 0; k < NU; ++k) pU[k] = okU[k] ? uPl[offU[k]] : 0.0;
#pragma unroll
    for (int k = 0; k < NV; ++k) pV[k] = okV[k] ? vPl[offV[k]] : 0.0;
    if (col) {
        pcs = *sCol;
        pw = wCol[wks];
        pf = *fCol;
    }

    for (int q = z0 - 1; q <= z1 + 1; ++q) {
        const bool zface = q >= z0;        // z-face q is in the chunk's range
        const bool pplane = q >= z0 + 1;   // plane p = q - 1 is a chunk plane
        const bool first = q == z0 - 1;
        double* ULq = UL[slot_q];
        double* VLq = VL[slot_q];
        const double* ULp = UL[slot_p];
        const double* VLp = VL[slot_p];

        // ---- A: stage plane q (loaded during the previous step); shift the
        //      column window; issue the loads of plane q + 1
#pragma unroll
        for (int k = 0; k < NS; ++k)
            if (okS[k]) S[tid + k * NT] = pS[k];
#pragma unroll
        for (int k = 0; k < NU; ++k)
            if (okU[k]) {
                ULq[tid + k * NT] = pU[k];
                SGX[tid + k * NT] = dabs(pU[k]) * dt / hx;
            }
#pragma unroll
        for (int k = 0; k < NV; ++k)
            if (okV[k]) {
                VLq[tid + k * NT] = pV[k];
                SGY[tid + k * NT] = dabs(pV[k]) * dt / hy;
            }
        if (col) {
#pragma unroll
            for (int jj = 0; jj < 8; ++jj) v[jj] = v[jj + 1];
            v[4 + G] = pcs;
            w_hi = pw;
            fo_q = pf;
        }
        if (q <= z1) {
            sPl += sks;
            uPl += uks;
            vPl += vks;
#pragma unroll
            for (int k = 0; k < NS; ++k)
                if (okS[k]) pS[k] = sPl[offS[k]];
#pragma unroll
            for (int k = 0; k < NU; ++k)
                if (okU[k]) pU[k] = uPl[offU[k]];
#pragma unroll
            for (int k = 0; k < NV; ++k)
                if (okV[k]) pV[k] = vPl[offV[k]];
            if (col) {
                sCol += sks;
                fCol += fks;
                pcs = *sCol;
                pw = wCol[2 * wks];
                pf = *fCol;
            }
        }
        __syncthreads();

        // ---- B1: PPM face values of plane q; PPM of cell q in z
        {
            double* FX = P1;
            double* FY = P1 + FXW * H;
            for (int idx = tid; idx < FXW * H; idx += NT) {
                const int f = idx % FXW - FO, r = idx / FXW;  // face f: cells f-1 | f
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
                const int c = idx % W, f = idx / W - FO;  // face f: cells f-1 | f
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
        }
        double Ipz = 0.0, Imz = 0.0;
        if (col) {
            SC[tid] = v[3];
            FL[tid] = fo_p;
            double sm, sp;
            if (PPM == 1) {
                // zf0 = face q, zsl = slope(q); new: slope(q+1), face q+1
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
                // zf0, zf1, zf2 = faces q-1, q, q+1; new: face q+2
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
            const double sig_hi = dabs(w_hi) * dt / hz;
            trace_sig(v[4], sm, sp, w_hi, w_q, sig_hi, sig_lo, rel_eps, Ipz, Imz);
            sig_lo = sig_hi;
            if (zface) {
                slz = Ipz_prev;
                srz = Imz;
                bc_pair(q, dlz, dhz, bcz0, bcz1, vn2, v[3], v[4], slz, srz);
                hz_q = upwind(slz, srz, w_q, 0.0);
            }
            Ipz_prev = Ipz;
        }
        __syncthreads();

        // ---- B2: PPM of plane q in x and y
        if (col) {
            const double* FX = P1;
            const double* FY = P1 + FXW * H;
            double sm, sp, Ip, Im;
            double w9[9];
#pragma unroll
            for (int o = -4; o <= 4; ++o) w9[4 + o] = (o >= -G && o <= G) ? S[si + o] : 0.0;
            const int fx = (a + FO) + b * FXW;
            if (PPM == 1) {
                sm = FX[fx];
                sp = FX[fx + 1];
                cw_limit(w9[4], sm, sp);
            } else {
                cs_cell(w9, FX[fx - 1], FX[fx], FX[fx + 1], FX[fx + 2], 1.e-10, sm, sp);
            }
            if (!XYF && ppm_special<PPM>(I, dlx, dhx, bcx0, bcx1))
                ppm_edges<PPM, 0>(w9, I, dlx, dhx, bcx0, bcx1, sm, sp);
            trace_sig(w9[4], sm, sp, ULq[ui + 1], ULq[ui], SGX[ui + 1], SGX[ui], rel_eps, Ip,
                      Im);
            P2[tid] = Ip;
            P2[NE + tid] = Im;
#pragma unroll
            for (int o = -4; o <= 4; ++o) w9[4 + o] = (o >= -G && o <= G) ? S[si + o * SW] : 0.0;
            const int fy = a + (b + FO) * W;
            if (PPM == 1) {
                sm = FY[fy];
                sp = FY[fy + W];
                cw_limit(w9[4], sm, sp);
            } else {
                cs_cell(w9, FY[fy - W], FY[fy], FY[fy + W], FY[fy + 2 * W], 1.e-10, sm, sp);
            }
            if (!XYF && ppm_special<PPM>(J, dly, dhy, bcy0, bcy1))
                ppm_edges<PPM, 1>(w9, J, dly, dhy, bcy0, bcy1, sm, sp);
            trace_sig(w9[4], sm, sp, VLq[vi + W], VLq[vi], SGY[vi + W], SGY[vi], rel_eps, Ip,
                      Im);
            P2[2 * NE + tid] = Ip;
            P2[3 * NE + tid] = Im;
        }
        __syncthreads();

        // ---- C: predictor and upwind states on the x- and y-faces of plane q
        if (col && a >= 1) {
            slx = P2[tid - 1];
            srx = P2[NE + tid];
            if (!XYF) bc_pair(I, dlx, dhx, bcx0, bcx1, vn0, S[si - 1], S[si], slx, srx);
            P1[tid] = upwind(slx, srx, ULq[ui], 0.0);
        }
        if (col && b >= 1) {
            sly = P2[2 * NE + tid - W];
            sry = P2[3 * NE + tid];
            if (!XYF) bc_pair(J, dly, dhy, bcy0, bcy1, vn1, S[si - SW], S[si], sly, sry);
            P1[NE + tid] = upwind(sly, sry, VLq[vi], 0.0);
        }
        __syncthreads();

        // ---- D1: per-cell transverse terms
        if (col) {
            const double ux1 = ULq[ui + 1], ux0 = ULq[ui];
            const double vy1 = VLq[vi + W], vy0 = VLq[vi];
            if (CONS) {
                double dv = ux1 - ux0 + vy1 - vy0 + w_hi - w_q;
                dv /= hx;
                bq = dt3 * v[4] * dv;
                P2[6 * NE + tid] = bq;
                P2[7 * NE + tid] = bp;
            }
            if (a >= 1 && a <= nx) {
                const double h1 = P1[tid + 1], h0 = P1[tid];
                if (CONS) {
                    tx_q = tfx * (h1 * ux1 - h0 * ux0);
                    cx_q = tfx * v[4] * (ux1 - ux0);
                    P2[3 * NE + tid] = cx_q;
                } else {
                    tx_q = tfx * (ux1 + ux0) * (h1 - h0);
                }
                P2[tid] = tx_q;
            }
            if (b >= 1 && b <= ny) {
                const double h1 = P1[NE + tid + W], h0 = P1[NE + tid];
                if (CONS) {
                    ty_q = tfy * (h1 * vy1 - h0 * vy0);
                    cy_q = tfy * v[4] * (vy1 - vy0);
                    P2[4 * NE + tid] = cy_q;
                } else {
                    ty_q = tfy * (vy1 + vy0) * (h1 - h0);
                }
                P2[NE + tid] = ty_q;
            }
            if (pplane) {
                if (CONS) {
                    P2[2 * NE + tid] = tfz * (hz_q * w_q - hz_p * w_p);
                    P2[5 * NE + tid] = tfz * v[3] * (w_q - w_p);
                } else {
                    P2[2 * NE + tid] = tfz * (w_q + w_p) * (hz_q - hz_p);
                }
            }
        }
        __syncthreads();

        // ---- D2: corner-coupled states
        if (col && a >= 1 && b >= 1 && b <= ny) {  // simhxy, x-faces of plane q
            double l = apply3<CONS>(slx, P2[NE + tid - 1], CONS ? P2[6 * NE + tid - 1] : 0.0,
                                    CONS ? P2[4 * NE + tid - 1] : 0.0);
            double r = apply3<CONS>(srx, P2[NE + tid], CONS ? P2[6 * NE + tid] : 0.0,
                                    CONS ? P2[4 * NE + tid] : 0.0);
            if (!XYF) bc_pair(I, dlx, dhx, bcx0, bcx1, vn0, S[si - 1], S[si], l, r);
            P1[tid] = upwind(l, r, ULq[ui], rel_eps);
        }
        if (col && b >= 1 && a >= 1 && a <= nx) {  // simhyx, y-faces of plane q
            double l = apply3<CONS>(sly, P2[tid - W], CONS ? P2[6 * NE + tid - W] : 0.0,
                                    CONS ? P2[3 * NE + tid - W] : 0.0);
            double r = apply3<CONS>(sry, P2[tid], CONS ? P2[6 * NE + tid] : 0.0,
                                    CONS ? P2[3 * NE + tid] : 0.0);
            if (!XYF) bc_pair(J, dly, dhy, bcy0, bcy1, vn1, S[si - SW], S[si], l, r);
            P1[NE + tid] = upwind(l, r, VLq[vi], rel_eps);
        }
        if (pplane && col && a >= 1) {  // simhxz, x-faces of plane p
            double l = apply3<CONS>(slx_p, P2[2 * NE + tid - 1], CONS ? P2[7 * NE + tid - 1] : 0.0,
                                    CONS ? P2[5 * NE + tid - 1] : 0.0);
            double r = apply3<CONS>(srx_p, P2[2 * NE + tid], CONS ? P2[7 * NE + tid] : 0.0,
                                    CONS ? P2[5 * NE + tid] : 0.0);
            if (!XYF) bc_pair(I, dlx, dhx, bcx0, bcx1, vn0, SC[tid - 1], SC[tid], l, r);
            P1[2 * NE + tid] = upwind(l, r, ULp[ui], rel_eps);
        }
        if (pplane && col && b >= 1) {  // simhyz, y-faces of plane p
            double l = apply3<CONS>(sly_p, P2[2 * NE + tid - W], CONS ? P2[7 * NE + tid - W] : 0.0,
                                    CONS ? P2[5 * NE + tid - W] : 0.0);
            double r = apply3<CONS>(sry_p, P2[2 * NE + tid], CONS ? P2[7 * NE + tid] : 0.0,
                                    CONS ? P2[5 * NE + tid] : 0.0);
            if (!XYF) bc_pair(J, dly, dhy, bcy0, bcy1, vn1, SC[tid - W], SC[tid], l, r);
            P1[3 * NE + tid] = upwind(l, r, VLp[vi], rel_eps);
        }
        if (zface && col && a >= 1 && a <= nx) {  // simhzx, z-face q
            double l = apply3<CONS>(slz, tx_p, bp, cx_p);
            double r = apply3<CONS>(srz, tx_q, bq, cx_q);
            bc_pair(q, dlz, dhz, bcz0, bcz1, vn2, v[3], v[4], l, r);
            tzx_q = upwind(l, r, w_q, rel_eps);
        }
        if (zface && col && b >= 1 && b <= ny) {  // simhzy, z-face q
            double l = apply3<CONS>(slz, ty_p, bp, cy_p);
            double r = apply3<CONS>(srz, ty_q, bq, cy_q);
            bc_pair(q, dlz, dhz, bcz0, bcz1, vn2, v[3], v[4], l, r);
            tzy_q = upwind(l, r, w_q, rel_eps);
        }
        __syncthreads();

        // ---- E1: per-cell edge terms
        if (col && a >= 1 && a <= nx && b >= 1 && b <= ny) {
            const double ux1 = ULq[ui + 1], ux0 = ULq[ui];
            const double vy1 = VLq[vi + W], vy0 = VLq[vi];
            const double t1 = P1[tid + 1], t0 = P1[tid];
            const double y1 = P1[NE + tid + W], y0 = P1[NE + tid];
            if (CONS) {
                exy_q = efx * (t1 * ux1 - t0 * ux0);
                eyx_q = efy * (y1 * vy1 - y0 * vy0);
            } else {
                exy_q = efx * (ux1 + ux0) * (t1 - t0);
                eyx_q = efy * (vy1 + vy0) * (y1 - y0);
            }
        }
        if (pplane && col) {
            if (b >= 1 && b <= ny) {
                const double vy1 = VLp[vi + W], vy0 = VLp[vi];
                const double t1 = P1[3 * NE + tid + W], t0 = P1[3 * NE + tid];
                P2[tid] = CONS ? efy * (t1 * vy1 - t0 * vy0) : efy * (vy1 + vy0) * (t1 - t0);
                P2[NE + tid] = CONS ? efz * (tzy_q * w_q - tzy_p * w_p)
                                    : efz * (w_q + w_p) * (tzy_q - tzy_p);
            }
            if (a >= 1 && a <= nx) {
                const double ux1 = ULp[ui + 1], ux0 = ULp[ui];
                const double t1 = P1[2 * NE + tid + 1], t0 = P1[2 * NE + tid];
                P2[2 * NE + tid] = CONS ? efx * (t1 * ux1 - t0 * ux0) : efx * (ux1 + ux0) * (t1 - t0);
                P2[3 * NE + tid] = CONS ? efz * (tzx_q * w_q - tzx_p * w_p)
                                        : efz * (w_q + w_p) * (tzx_q - tzx_p);
            }
        }
        __syncthreads();

        // ---- E2: edge states
        if (pplane && col && a >= 1 && b >= 1 && b <= ny && (a <= nx || last_x)) {  // x-faces, p
            double l = slx_p - P2[tid - 1];
            l = l - P2[NE + tid - 1];
            double r = srx_p - P2[tid];
            r = r - P2[NE + tid];
            const double u0 = ULp[ui];
            if (CONS) {
                l = l - dt2hx * SC[tid - 1] * (u0 - ULp[ui - 1]);
                r = r - dt2hx * SC[tid] * (ULp[ui + 1] - u0);
            }
            const double sedgel = l + dt2 * FL[tid - 1];
            const double sedger = r + dt2 * FL[tid];
            double res = upwind(sedgel, sedger, u0, rel_eps);
            if (!XYF)
                res = bc_edge(res, I, dlx, dhx, bcx0, bcx1, vn0, SC[tid - 1], SC[tid], sedgel,
                              sedger);
            e0Col[-e0ks] = res;  // plane p
            if (flux) flux_store(boxes[bid], 0, I, J, q - 1, comp, flux, u0, res, false);
        }
        if (pplane && col && b >= 1 && a >= 1 && a <= nx && (b <= ny || last_y)) {  // y-faces, p
            double l = sly_p - P2[2 * NE + tid - W];
            l = l - P2[3 * NE + tid - W];
            double r = sry_p - P2[2 * NE + tid];
            r = r - P2[3 * NE + tid];
            const double u0 = VLp[vi];
            if (CONS) {
                l = l - dt2hy * SC[tid - W] * (u0 - VLp[vi - W]);
                r = r - dt2hy * SC[tid] * (VLp[vi + W] - u0);
            }
            const double sedgel = l + dt2 * FL[tid - W];
            const double sedger = r + dt2 * FL[tid];
            double res = upwind(sedgel, sedger, u0, rel_eps);
            if (!XYF)
                res = bc_edge(res, J, dly, dhy, bcy0, bcy1, vn1, SC[tid - W], SC[tid], sedgel,
                              sedger);
            e1Col[-e1ks] = res;  // plane p
            if (flux) flux_store(boxes[bid], 1, I, J, q - 1, comp, flux, u0, res, false);
        }
        if (((zface && q <= z1) || (q == z1 + 1 && last_z)) && col && a >= 1 && a <= nx &&
            b >= 1 && b <= ny) {  // z-face q
            double l = slz - exy_p;
            l = l - eyx_p;
            double r = srz - exy_q;
            r = r - eyx_q;
            if (CONS) {
                l = l - dt2hz * v[3] * (w_q - w_p);
                r = r - dt2hz * v[4] * (w_hi - w_q);
            }
            const double sedgel = l + dt2 * fo_p;
            const double sedger = r + dt2 * fo_q;
            double res = upwind(sedgel, sedger, w_q, rel_eps);
            res = bc_edge(res, q, dlz, dhz, bcz0, bcz1, vn2, v[3], v[4], sedgel, sedger);
            *e2Col = res;  // z-face q
            if (flux) flux_store(boxes[bid], 2, I, J, q, comp, flux, w_q, res, true);
        }

        // ---- advance the column state to the next plane
        slx_p = slx;
        srx_p = srx;
        sly_p = sly;
        sry_p = sry;
        hz_p = hz_q;
        tx_p = tx_q;
        ty_p = ty_q;
        cx_p = cx_q;
        cy_p = cy_q;
        bp = bq;
        tzx_p = tzx_q;
        tzy_p = tzy_q;
        exy_p = exy_q;
        eyx_p = eyx_q;
        w_p = w_q;
        w_q = w_hi;
        fo_p = fo_q;
        wCol += wks;
        e0Col += e0ks;
        e1Col += e1ks;
        e2Col += e2ks;
        slot_p = slot_q;
        slot_q = slot_q == 2 ? 0 : slot_q + 1;
    }
}

}  // namespace

namespace mxh {

bool fused_edge_scal_ok(const mx_adv_ctx* c, int is_conservative) {
    (void)is_conservative;
    return c->p.dim == 3 && (c->p.ppm_type == 1 || c->p.ppm_type == 2) &&
           c->p.ppm_trace_forces == 0 && c->fused_edge;
}

namespace {

// does boundary code `bc` override anything (everything but INT_DIR)?
bool bc_acts(int bc) {
    return bc == MX_BC_EXT_DIR || bc == MX_BC_HOEXTRAP || bc == MX_BC_FOEXTRAP ||
           bc == MX_BC_REFLECT_EVEN || bc == MX_BC_REFLECT_ODD;
}

// A tile is x/y-free when, in x and in y, each domain side either carries no
// override for any component or lies beyond the tile's reach: the kernel
// evaluates cells [X0 - 1, X0 + nx] (PPM boundary stencils act within three
// cells of a side) and faces [X0, X0 + nx] (face overrides act on the side's
// face).
bool tile_xy_free(const int lo0, int n, int dlo, int dhi, bool act_lo, bool act_hi) {
    const bool free_lo = !act_lo || lo0 - 1 > dlo + 2;
    const bool free_hi = !act_hi || lo0 + n < dhi - 2;
    return free_lo && free_hi;
}

template <int TX, int TY, int WPE, bool XYF>
int launch_one(mx_adv_ctx* c, const FBox* db, const FTile* dtl, size_t ntile, const FComp* dc,
               const KGeom& kg, const FArgs& A, int num_comp, bool cons) {
    if (ntile == 0) return MX_OK;
    const int xm = c->es_fused_xcd;
    const dim3 grid = xm ? dim3(xcd_tile_blocks((long long)ntile, num_comp))
                         : dim3((unsigned)ntile, 1, (unsigned)num_comp);
    const int nt = (int)ntile;
    dim3 block(fused_nt(TX, TY));
    if (c->p.ppm_type == 1) {
        if (cons)
            MX_LAUNCH(c, "k_es_fused", (k_es_fused<1, true, TX, TY, XYF, WPE>), grid, block, db, dtl, dc, kg, A, nt, xm);
        else
            MX_LAUNCH(c, "k_es_fused", (k_es_fused<1, false, TX, TY, XYF, WPE>), grid, block, db, dtl, dc, kg, A, nt, xm);
    } else {
        if (cons)
            MX_LAUNCH(c, "k_es_fused", (k_es_fused<2, true, TX, TY, XYF, WPE>), grid, block, db, dtl, dc, kg, A, nt, xm);
        else
            MX_LAUNCH(c, "k_es_fused", (k_es_fused<2, false, TX, TY, XYF, WPE>), grid, block, db, dtl, dc, kg, A, nt, xm);
    }
    return MX_OK;
}

template <int TX, int TY, int WPE>
int launch_fused(mx_adv_ctx* c, const std::vector<EdgeLevel>& levs, int tz, int is_vel,
                 const mx_bcrec* bcs, const std::vector<int>& scomps,
                 const std::vector<int>& bccomps, int is_conservative) {
    const int num_comp = (int)scomps.size();
    const long long lim = 0x7fffffffLL;
    std::vector<FBox> boxes;
    std::vector<FTile> tiles, btiles;  // x/y-free tiles, then the others
    bool act[2][2] = {{false, false}, {false, false}};
    for (int vc = 0; vc < num_comp; ++vc)
        for (int d = 0; d < 2; ++d) {
            act[d][0] = act[d][0] || bc_acts(bcs[bccomps[vc]].lo[d]);
            act[d][1] = act[d][1] || bc_acts(bcs[bccomps[vc]].hi[d]);
        }
    for (const EdgeLevel& L : levs) {
        const auto& grids = c->grids[L.lev];
        for (size_t b = 0; b < grids.size(); ++b) {
            if (L.box >= 0 && (size_t)L.box != b) continue;
            FBox B;
            HBox g = hbox(grids[b]);
            for (int q = 0; q < 3; ++q) {
                B.lo[q] = g.lo[q];
                B.hi[q] = g.hi[q];
                B.dx[q] = c->p.dx[L.lev][q];
                B.domlo[q] = c->p.domain_lo[L.lev][q];
                B.domhi[q] = c->p.domain_hi[L.lev][q];
            }
            B.s = to_da(L.state->fab[b]);
            B.fo = to_da(L.force->fab[b]);
            for (int d = 0; d < 3; ++d) {
                B.um[d] = to_da(L.umac->dir[d].fab[b]);
                B.se[d] = to_da(L.sedge->dir[d].fab[b]);
                B.sf[d] = L.sflux ? to_da(L.sflux->dir[d].fab[b]) : null_da();
            }
            B.bc_a = L.bc_a;
            B.bc_b = L.bc_b;
            B.be_a = L.be_a;
            B.be_b = L.be_b;
            // the kernel keeps the plane strides in 32 bits
            const DA* views[] = {&B.s,     &B.fo,    &B.um[0], &B.um[1],
                                 &B.um[2], &B.se[0], &B.se[1], &B.se[2]};
            for (const DA* d : views)
                if (d->js > lim || d->ks > lim) return MX_ERR_UNSUPPORTED;
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
                            tile_xy_free(x0, nx, B.domlo[0], B.domhi[0], act[0][0], act[0][1]) &&
                            tile_xy_free(y0, ny, B.domlo[1], B.domhi[1], act[1][0], act[1][1]);
                        (f ? tiles : btiles).push_back({id, x, y, z});
                    }
            boxes.push_back(B);
        }
    }
    const size_t nfree = tiles.size();
    tiles.insert(tiles.end(), btiles.begin(), btiles.end());
    if (tiles.empty()) return MX_OK;
    if (tiles.size() > 0x7fffffffULL) return MX_ERR_UNSUPPORTED;
    std::vector<FComp> comps(num_comp);
    for (int vc = 0; vc < num_comp; ++vc) {
        for (int d = 0; d < 3; ++d) {
            comps[vc].bclo[d] = bcs[bccomps[vc]].lo[d];
            comps[vc].bchi[d] = bcs[bccomps[vc]].hi[d];
        }
        comps[vc].comp = scomps[vc];
        comps[vc].flux = (levs[0].sflux && scomps[vc] == levs[0].flux_comp) ? levs[0].flux_kind : 0;
    }
    FArgs A;
    A.start_scomp = scomps[0];
    A.num_comp = num_comp;
    A.is_vel = is_vel;
    A.tz = tz;
    Table tab(c);
    size_t ob = tab.add(boxes);
    size_t ot = tab.add(tiles);
    size_t oc = tab.add(comps);
    int rc = tab.upload();
    if (rc) return rc;
    const FBox* db = tab.dev<FBox>(ob);
    const FTile* dtl = tab.dev<FTile>(ot);
    const FComp* dc = tab.dev<FComp>(oc);
    KGeom kg = kgeom(c, levs[0].lev);
    const bool cons = is_conservative != 0;
    rc = launch_one<TX, TY, WPE, true>(c, db, dtl, nfree, dc, kg, A, num_comp, cons);
    if (rc) return rc;
    return launch_one<TX, TY, WPE, false>(c, db, dtl + nfree, tiles.size() - nfree, dc, kg, A,
                                     num_comp, cons);
}

// the tile shape of `fused_tile`, with the block size (columns rounded up to
// whole waves) and the register budget (waves per SIMD):
//   0: 32 x 8 (384 threads), 1: 16 x 16 (384), 2: 16 x 12 (256), 3: 16 x 8 (192),
//   4: 16 x 11 (256)
// all at 3 waves per SIMD (168 VGPRs), so a CU holds 2, 2, 3, 4 or 3 blocks.
// -1 (the default) is 16 x 11: with ppm_type 1 its face phases (19 x 13
// x-faces, 18 x 14 y-faces) fit one pass of 256 threads, where 16 x 12
// leaves one wave a second pass while the others wait at the barrier; with
// ppm_type 2 its x/y-free variant spills 1 VGPR against 7 for 16 x 12
// (5.50 against 5.57 ms per 256^3 pass)
int fused_tile_of(const mx_adv_ctx* c) {
    if (c->fused_tile >= 0) return c->fused_tile;
    return 4;
}
int launch_tile(mx_adv_ctx* c, const std::vector<EdgeLevel>& levs, int tz, int is_vel,
                const mx_bcrec* bcs, const std::vector<int>& sc, const std::vector<int>& bc,
                int cons) {
    switch (fused_tile_of(c)) {
        case 1: return launch_fused<16, 16, 3>(c, levs, tz, is_vel, bcs, sc, bc, cons);
        case 2: return launch_fused<16, 12, 3>(c, levs, tz, is_vel, bcs, sc, bc, cons);
        case 3: return launch_fused<16, 8, 3>(c, levs, tz, is_vel, bcs, sc, bc, cons);
        case 4: return launch_fused<16, 11, 3>(c, levs, tz, is_vel, bcs, sc, bc, cons);
        default: return launch_fused<32, 8, 3>(c, levs, tz, is_vel, bcs, sc, bc, cons);
    }
}

}  // namespace

int make_edge_scal_fused_list(mx_adv_ctx* c, const std::vector<EdgeLevel>& levs, int is_vel,
                              const mx_bcrec* bcs, const std::vector<int>& scomps,
                              const std::vector<int>& bccomps, int is_conservative) {
    if (levs.empty() || scomps.empty()) return MX_OK;
    if (scomps.size() != bccomps.size()) return MX_ERR_ARG;
    const int tz = c->fused_tz > 0 ? c->fused_tz : 32;
    if (c->fused_per_box) {
        // comparison mode: the same work as one launch per box
        std::vector<EdgeLevel> one(1);
        for (const EdgeLevel& L : levs)
            for (size_t b = 0; b < c->grids[L.lev].size(); ++b) {
                one[0] = L;
                one[0].box = (int)b;
                int rc = launch_tile(c, one, tz, is_vel, bcs, scomps, bccomps, is_conservative);
                if (rc) return rc;
            }
        return MX_OK;
    }
    return launch_tile(c, levs, tz, is_vel, bcs, scomps, bccomps, is_conservative);
}

int make_edge_scal_fused_levels(mx_adv_ctx* c, const std::vector<EdgeLevel>& levs, int is_vel,
                                const mx_bcrec* bcs, int start_scomp, int start_bccomp,
                                int num_comp, int is_conservative) {
    std::vector<int> sc(num_comp), bc(num_comp);
    for (int vc = 0; vc < num_comp; ++vc) {
        sc[vc] = start_scomp + vc;
        bc[vc] = start_bccomp + vc;
    }
    return make_edge_scal_fused_list(c, levs, is_vel, bcs, sc, bc, is_conservative);
}

int make_edge_scal_fused(mx_adv_ctx* c, int lev, const mx_multifab* state,
                         const mx_face_multifab* sedge, const mx_face_multifab* umac,
                         const mx_multifab* force, int is_vel, const mx_bcrec* bcs,
                         int start_scomp, int start_bccomp, int num_comp, int is_conservative) {
    std::vector<EdgeLevel> levs(1);
    levs[0] = {lev, state, sedge, umac, force};
    return make_edge_scal_fused_levels(c, levs, is_vel, bcs, start_scomp, start_bccomp, num_comp,
                                       is_conservative);
}

}  // namespace mxh
