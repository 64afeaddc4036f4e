// mx_edge_fused.hip -- fused 3-D MakeEdgeScal: PPM, predictor, transverse
// and edge states of one scalar in a single kernel.
//
// A block owns a TX x TY tile of cell columns of one box and marches up a
// chunk [z0, z1] of its planes.  Every per-plane quantity lives in LDS on the
// tile grown by one cell in x and y (E = (TX+2) x (TY+2) points); at step q:
//   A  the scalar plane q (with the PPM stencil halo) is staged in LDS; each
//      column owner shifts its register window of s along z;
//   B  PPM in z of cell q (register window) and in x, y of plane q (LDS
//      stencils), traced with the MAC velocities;
//   C  predictor states slx/srx, sly/sry of plane q with their boundary
//      overrides and the upwind states simhx, simhy; slz/srz and simhz of the
//      z-face q were formed in B;
//   D  the six corner-coupled states: simhxy, simhyx on plane q, simhzx,
//      simhzy on z-face q, simhxz, simhyz on plane q-1;
//   E  the edge states of plane q-1 (x and y faces) and of z-face q.
// Ring buffers keep the two newest planes of the quantities read across
// planes; the predictor states a thread reads only at its own face stay in
// its registers.  The arithmetic of every point is that of
// MaestroMakeEdgeScalars.cpp (Predictor :662-863, Transverse :865-1359,
// Edges :1361-1667) and MaestroPPM.cpp, in the same operand order, so the
// results are bit-identical to the per-stage kernels and to the CPU path.
#include <algorithm>

#include "mx_internal.H"

using namespace mxd;
using namespace mxh;

namespace {

struct FBox {
    int lo[3], hi[3];
    DA s, fo;
    DA um[3], se[3];
};

struct FComp {
    int bclo[3], bchi[3];
};

struct FArgs {
    int start_scomp;
    int num_comp;
    int is_vel;
    int is_conservative;
    int tz;  // planes per chunk
};

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

__device__ __forceinline__ bool bc_needs_s(int pos, int dlo, int dhi, int bclo, int bchi) {
    return (pos == dlo && bclo == MX_BC_EXT_DIR) || (pos == dhi + 1 && bchi == MX_BC_EXT_DIR);
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

__device__ __forceinline__ double divu_at(const FBox& B, double dx0, int i, int j, int k) {
    double dv = B.um[0](i + 1, j, k) - B.um[0](i, j, k) + B.um[1](i, j + 1, k) -
                B.um[1](i, j, k) + B.um[2](i, j, k + 1) - B.um[2](i, j, k);
    dv /= dx0;
    return dv;
}

// corner-coupled update of one side (MakeEdgeScalTransverse): u1/u0 the
// transverse MAC velocities at the +1 and 0 faces, h1/h0 the upwind states
template <bool CONS>
__device__ __forceinline__ double trans_side(double x, double dt3h, double dt6h, double dt3,
                                             double u1, double u0, double h1, double h0, double s,
                                             double dv) {
    if (CONS) {
        return x - dt3h * (h1 * u1 - h0 * u0) - dt3 * s * dv + dt3h * s * (u1 - u0);
    } else {
        return x - dt6h * (u1 + u0) * (h1 - h0);
    }
}

template <bool CONS>
__device__ __forceinline__ double edge_side(double x, double dt2h, double dt4h, double u1,
                                            double u0, double h1, double h0) {
    if (CONS) {
        return x - dt2h * (h1 * u1 - h0 * u0);
    } else {
        return x - dt4h * (u1 + u0) * (h1 - h0);
    }
}

template <int PPM, int TX, int TY, bool CONS>
__global__ void __launch_bounds__(384) k_es_fused(const FBox* __restrict__ boxes,
                                                  const FComp* __restrict__ comps, KGeom g,
                                                  FArgs A) {
    constexpr int G = PPM == 2 ? 3 : 2;
    constexpr int W = TX + 2;
    constexpr int NE = (TX + 2) * (TY + 2);
    constexpr int SW = TX + 2 + 2 * G;
    constexpr int SH = TY + 2 + 2 * G;
    constexpr int NT = 384;
    static_assert(NE <= NT, "tile too large for the block");

    __shared__ double S[SW * SH];
    __shared__ double PX[NE], MX[NE], PY[NE], MY[NE];
    __shared__ double HX[2][NE], HY[2][NE], HZ[2][NE];
    __shared__ double TXY[2][NE], TYX[2][NE], TZX[2][NE], TZY[2][NE];
    __shared__ double TXZ[NE], TYZ[NE];

    const FBox& B = boxes[blockIdx.y];
    const int vc = blockIdx.z;
    const FComp& Cb = comps[vc];
    const int comp = A.start_scomp + vc;

    const int nbx = B.hi[0] - B.lo[0] + 1;
    const int nby = B.hi[1] - B.lo[1] + 1;
    const int nbz = B.hi[2] - B.lo[2] + 1;
    const int ntx = (nbx + TX - 1) / TX;
    const int nty = (nby + TY - 1) / TY;
    const int nzc = (nbz + A.tz - 1) / A.tz;
    const int tile = blockIdx.x;
    if (tile >= ntx * nty * nzc) return;
    const int tx = tile % ntx;
    const int ty = (tile / ntx) % nty;
    const int tzc = tile / (ntx * nty);
    const int X0 = B.lo[0] + tx * TX;
    const int Y0 = B.lo[1] + ty * TY;
    const int nx = min(TX, B.hi[0] - X0 + 1);
    const int ny = min(TY, B.hi[1] - Y0 + 1);
    const int z0 = B.lo[2] + tzc * A.tz;
    const int z1 = min(z0 + A.tz - 1, B.hi[2]);
    const bool last_z = (z1 == B.hi[2]);
    const bool last_x = (X0 + nx - 1 == B.hi[0]);
    const bool last_y = (Y0 + ny - 1 == B.hi[1]);

    const int tid = threadIdx.x;
    const int a = tid % W;  // local cell x + 1
    const int b = tid / W;  // local cell y + 1
    const bool col = (tid < NE) && (a <= nx + 1) && (b <= ny + 1);
    const int I = X0 + a - 1;
    const int J = Y0 + b - 1;

    const double dt = g.dt;
    const double dt2 = 0.5 * dt;
    const double dt3 = dt / 3.0;
    const double dt4 = 0.25 * dt;
    const double dt6 = dt / 6.0;
    const double hx = g.dx[0], hy = g.dx[1], hz = g.dx[2];
    const double dt6hx = dt6 / hx, dt6hy = dt6 / hy, dt6hz = dt6 / hz;
    const double dt3hx = dt3 / hx, dt3hy = dt3 / hy, dt3hz = dt3 / hz;
    const double dt4hx = dt4 / hx, dt4hy = dt4 / hy, dt4hz = dt4 / hz;
    const double dt2hx = dt2 / hx, dt2hy = dt2 / hy, dt2hz = dt2 / hz;
    const bool vn0 = A.is_vel && comp == 0;
    const bool vn1 = A.is_vel && comp == 1;
    const bool vn2 = A.is_vel && comp == 2;
    const DA& s = B.s;
    const DA& U = B.um[0];
    const DA& V = B.um[1];
    const DA& Wm = B.um[2];

    // register state of the owned column / faces
    double v[9];
    double Ipz_prev = 0.0, slz = 0.0, srz = 0.0;
    double slx = 0.0, srx = 0.0, slx_p = 0.0, srx_p = 0.0;
    double sly = 0.0, sry = 0.0, sly_p = 0.0, sry_p = 0.0;

    // v[jj] = s(I, J, z0 - 6 + jj): after the first shift v[4] = s(z0 - 1)
    for (int jj = 0; jj < 9; ++jj) v[jj] = 0.0;
    if (col) {
        for (int jj = 5 - G; jj <= 4 + G; ++jj) v[jj] = s(I, J, z0 - 6 + jj, comp);
    }

    for (int q = z0 - 1; q <= z1 + 1; ++q) {
        const int cq = q & 1;
        const int cp = cq ^ 1;
        const int p = q - 1;

        // ---- A: stage plane q, shift the z window
        for (int idx = tid; idx < SW * SH; idx += NT) {
            const int sx = idx % SW, sy = idx / SW;
            const int lx = sx - 1 - G, ly = sy - 1 - G;
            if (lx <= nx + G && ly <= ny + G) S[idx] = s(X0 + lx, Y0 + ly, q, comp);
        }
        if (col) {
#pragma unroll
            for (int jj = 0; jj < 8; ++jj) v[jj] = v[jj + 1];
            v[4 + G] = s(I, J, q + G, comp);
        }
        __syncthreads();

        // ---- B: PPM in z (register window) and in x, y (LDS plane)
        if (col) {
            double sm, sp, Ip, Im;
            const double wlo = Wm(I, J, q), whi = Wm(I, J, q + 1);
            ppm_edges<PPM, 2>(v, q, g.domlo[2], g.domhi[2], Cb.bclo[2], Cb.bchi[2], sm, sp);
            ppm_trace(v[4], sm, sp, whi, wlo, dt, hz, g.rel_eps, Ip, Im);
            if (q >= z0) {
                slz = Ipz_prev;
                srz = Im;
                bc_pair(q, g.domlo[2], g.domhi[2], Cb.bclo[2], Cb.bchi[2], vn2, v[3], v[4], slz,
                        srz);
                HZ[cq][tid] = upwind(slz, srz, wlo, 0.0);
            }
            Ipz_prev = Ip;

            const int si = (a + G) + (b + G) * SW;
            double w9[9];
            w9[0] = w9[8] = 0.0;
#pragma unroll
            for (int o = -3; o <= 3; ++o) {
                if (o >= -G && o <= G) w9[4 + o] = S[si + o];
                else w9[4 + o] = 0.0;
            }
            ppm_edges<PPM, 0>(w9, I, g.domlo[0], g.domhi[0], Cb.bclo[0], Cb.bchi[0], sm, sp);
            ppm_trace(w9[4], sm, sp, U(I + 1, J, q), U(I, J, q), dt, hx, g.rel_eps, Ip, Im);
            PX[tid] = Ip;
            MX[tid] = Im;
#pragma unroll
            for (int o = -3; o <= 3; ++o) {
                if (o >= -G && o <= G) w9[4 + o] = S[si + o * SW];
                else w9[4 + o] = 0.0;
            }
            ppm_edges<PPM, 1>(w9, J, g.domlo[1], g.domhi[1], Cb.bclo[1], Cb.bchi[1], sm, sp);
            ppm_trace(w9[4], sm, sp, V(I, J + 1, q), V(I, J, q), dt, hy, g.rel_eps, Ip, Im);
            PY[tid] = Ip;
            MY[tid] = Im;
        }
        __syncthreads();

        // ---- C: predictor states and upwind states on x- and y-faces of plane q
        if (col && a >= 1) {
            const int si = (a + G) + (b + G) * SW;
            slx = PX[tid - 1];
            srx = MX[tid];
            bc_pair(I, g.domlo[0], g.domhi[0], Cb.bclo[0], Cb.bchi[0], vn0, S[si - 1], S[si], slx,
                    srx);
            HX[cq][tid] = upwind(slx, srx, U(I, J, q), 0.0);
        }
        if (col && b >= 1) {
            const int si = (a + G) + (b + G) * SW;
            sly = PY[tid - W];
            sry = MY[tid];
            bc_pair(J, g.domlo[1], g.domhi[1], Cb.bclo[1], Cb.bchi[1], vn1, S[si - SW], S[si],
                    sly, sry);
            HY[cq][tid] = upwind(sly, sry, V(I, J, q), 0.0);
        }
        __syncthreads();

        // ---- D: corner-coupled states
        // simhxy on x-faces of plane q
        if (col && a >= 1 && b >= 1 && b <= ny) {
            const int si = (a + G) + (b + G) * SW;
            double sL = 0.0, sR = 0.0, dvL = 0.0, dvR = 0.0;
            if (CONS) {
                sL = S[si - 1];
                sR = S[si];
                dvL = divu_at(B, hx, I - 1, J, q);
                dvR = divu_at(B, hx, I, J, q);
            }
            double l = trans_side<CONS>(slx, dt3hy, dt6hy, dt3, V(I - 1, J + 1, q), V(I - 1, J, q),
                                        HY[cq][tid - 1 + W], HY[cq][tid - 1], sL, dvL);
            double r = trans_side<CONS>(srx, dt3hy, dt6hy, dt3, V(I, J + 1, q), V(I, J, q),
                                        HY[cq][tid + W], HY[cq][tid], sR, dvR);
            bc_pair(I, g.domlo[0], g.domhi[0], Cb.bclo[0], Cb.bchi[0], vn0, S[si - 1], S[si], l, r);
            TXY[cq][tid] = upwind(l, r, U(I, J, q), g.rel_eps);
        }
        // simhyx on y-faces of plane q
        if (col && b >= 1 && a >= 1 && a <= nx) {
            const int si = (a + G) + (b + G) * SW;
            double sL = 0.0, sR = 0.0, dvL = 0.0, dvR = 0.0;
            if (CONS) {
                sL = S[si - SW];
                sR = S[si];
                dvL = divu_at(B, hx, I, J - 1, q);
                dvR = divu_at(B, hx, I, J, q);
            }
            double l = trans_side<CONS>(sly, dt3hx, dt6hx, dt3, U(I + 1, J - 1, q), U(I, J - 1, q),
                                        HX[cq][tid - W + 1], HX[cq][tid - W], sL, dvL);
            double r = trans_side<CONS>(sry, dt3hx, dt6hx, dt3, U(I + 1, J, q), U(I, J, q),
                                        HX[cq][tid + 1], HX[cq][tid], sR, dvR);
            bc_pair(J, g.domlo[1], g.domhi[1], Cb.bclo[1], Cb.bchi[1], vn1, S[si - SW], S[si], l,
                    r);
            TYX[cq][tid] = upwind(l, r, V(I, J, q), g.rel_eps);
        }
        if (q >= z0 + 1) {
            // simhxz on x-faces of plane p = q - 1
            if (col && a >= 1) {
                double sL = 0.0, sR = 0.0, dvL = 0.0, dvR = 0.0;
                if (CONS || bc_needs_s(I, g.domlo[0], g.domhi[0], Cb.bclo[0], Cb.bchi[0])) {
                    sL = s(I - 1, J, p, comp);
                    sR = s(I, J, p, comp);
                }
                if (CONS) {
                    dvL = divu_at(B, hx, I - 1, J, p);
                    dvR = divu_at(B, hx, I, J, p);
                }
                double l = trans_side<CONS>(slx_p, dt3hz, dt6hz, dt3, Wm(I - 1, J, q),
                                            Wm(I - 1, J, p), HZ[cq][tid - 1], HZ[cp][tid - 1], sL,
                                            dvL);
                double r = trans_side<CONS>(srx_p, dt3hz, dt6hz, dt3, Wm(I, J, q), Wm(I, J, p),
                                            HZ[cq][tid], HZ[cp][tid], sR, dvR);
                bc_pair(I, g.domlo[0], g.domhi[0], Cb.bclo[0], Cb.bchi[0], vn0, sL, sR, l, r);
                TXZ[tid] = upwind(l, r, U(I, J, p), g.rel_eps);
            }
            // simhyz on y-faces of plane p
            if (col && b >= 1) {
                double sL = 0.0, sR = 0.0, dvL = 0.0, dvR = 0.0;
                if (CONS || bc_needs_s(J, g.domlo[1], g.domhi[1], Cb.bclo[1], Cb.bchi[1])) {
                    sL = s(I, J - 1, p, comp);
                    sR = s(I, J, p, comp);
                }
                if (CONS) {
                    dvL = divu_at(B, hx, I, J - 1, p);
                    dvR = divu_at(B, hx, I, J, p);
                }
                double l = trans_side<CONS>(sly_p, dt3hz, dt6hz, dt3, Wm(I, J - 1, q),
                                            Wm(I, J - 1, p), HZ[cq][tid - W], HZ[cp][tid - W], sL,
                                            dvL);
                double r = trans_side<CONS>(sry_p, dt3hz, dt6hz, dt3, Wm(I, J, q), Wm(I, J, p),
                                            HZ[cq][tid], HZ[cp][tid], sR, dvR);
                bc_pair(J, g.domlo[1], g.domhi[1], Cb.bclo[1], Cb.bchi[1], vn1, sL, sR, l, r);
                TYZ[tid] = upwind(l, r, V(I, J, p), g.rel_eps);
            }
        }
        if (q >= z0) {
            double dvL = 0.0, dvR = 0.0;
            if (CONS && col) {
                dvL = divu_at(B, hx, I, J, q - 1);
                dvR = divu_at(B, hx, I, J, q);
            }
            // simhzx on z-face q
            if (col && a >= 1 && a <= nx) {
                double l = trans_side<CONS>(slz, dt3hx, dt6hx, dt3, U(I + 1, J, q - 1),
                                            U(I, J, q - 1), HX[cp][tid + 1], HX[cp][tid], v[3],
                                            dvL);
                double r = trans_side<CONS>(srz, dt3hx, dt6hx, dt3, U(I + 1, J, q), U(I, J, q),
                                            HX[cq][tid + 1], HX[cq][tid], v[4], dvR);
                bc_pair(q, g.domlo[2], g.domhi[2], Cb.bclo[2], Cb.bchi[2], vn2, v[3], v[4], l, r);
                TZX[cq][tid] = upwind(l, r, Wm(I, J, q), g.rel_eps);
            }
            // simhzy on z-face q
            if (col && b >= 1 && b <= ny) {
                double l = trans_side<CONS>(slz, dt3hy, dt6hy, dt3, V(I, J + 1, q - 1),
                                            V(I, J, q - 1), HY[cp][tid + W], HY[cp][tid], v[3],
                                            dvL);
                double r = trans_side<CONS>(srz, dt3hy, dt6hy, dt3, V(I, J + 1, q), V(I, J, q),
                                            HY[cq][tid + W], HY[cq][tid], v[4], dvR);
                bc_pair(q, g.domlo[2], g.domhi[2], Cb.bclo[2], Cb.bchi[2], vn2, v[3], v[4], l, r);
                TZY[cq][tid] = upwind(l, r, Wm(I, J, q), g.rel_eps);
            }
        }
        __syncthreads();

        // ---- E: edge states
        if (q >= z0 + 1) {
            // x-faces of plane p
            if (col && a >= 1 && b >= 1 && b <= ny && (a <= nx || last_x)) {
                double l = slx_p, r = srx_p;
                l = edge_side<CONS>(l, dt2hy, dt4hy, V(I - 1, J + 1, p), V(I - 1, J, p),
                                    TYZ[tid - 1 + W], TYZ[tid - 1]);
                r = edge_side<CONS>(r, dt2hy, dt4hy, V(I, J + 1, p), V(I, J, p), TYZ[tid + W],
                                    TYZ[tid]);
                l = edge_side<CONS>(l, dt2hz, dt4hz, Wm(I - 1, J, q), Wm(I - 1, J, p),
                                    TZY[cq][tid - 1], TZY[cp][tid - 1]);
                r = edge_side<CONS>(r, dt2hz, dt4hz, Wm(I, J, q), Wm(I, J, p), TZY[cq][tid],
                                    TZY[cp][tid]);
                double sL = 0.0, sR = 0.0;
                if (CONS || bc_needs_s(I, g.domlo[0], g.domhi[0], Cb.bclo[0], Cb.bchi[0])) {
                    sL = s(I - 1, J, p, comp);
                    sR = s(I, J, p, comp);
                }
                if (CONS) {
                    l = l - dt2hx * sL * (U(I, J, p) - U(I - 1, J, p));
                    r = r - dt2hx * sR * (U(I + 1, J, p) - U(I, J, p));
                }
                const double sedgel = l + dt2 * B.fo(I - 1, J, p, comp);
                const double sedger = r + dt2 * B.fo(I, J, p, comp);
                double res = upwind(sedgel, sedger, U(I, J, p), g.rel_eps);
                res = bc_edge(res, I, g.domlo[0], g.domhi[0], Cb.bclo[0], Cb.bchi[0], vn0, sL, sR,
                              sedgel, sedger);
                B.se[0](I, J, p, comp) = res;
            }
            // y-faces of plane p
            if (col && b >= 1 && a >= 1 && a <= nx && (b <= ny || last_y)) {
                double l = sly_p, r = sry_p;
                l = edge_side<CONS>(l, dt2hx, dt4hx, U(I + 1, J - 1, p), U(I, J - 1, p),
                                    TXZ[tid - W + 1], TXZ[tid - W]);
                r = edge_side<CONS>(r, dt2hx, dt4hx, U(I + 1, J, p), U(I, J, p), TXZ[tid + 1],
                                    TXZ[tid]);
                l = edge_side<CONS>(l, dt2hz, dt4hz, Wm(I, J - 1, q), Wm(I, J - 1, p),
                                    TZX[cq][tid - W], TZX[cp][tid - W]);
                r = edge_side<CONS>(r, dt2hz, dt4hz, Wm(I, J, q), Wm(I, J, p), TZX[cq][tid],
                                    TZX[cp][tid]);
                double sL = 0.0, sR = 0.0;
                if (CONS || bc_needs_s(J, g.domlo[1], g.domhi[1], Cb.bclo[1], Cb.bchi[1])) {
                    sL = s(I, J - 1, p, comp);
                    sR = s(I, J, p, comp);
                }
                if (CONS) {
                    l = l - dt2hy * sL * (V(I, J, p) - V(I, J - 1, p));
                    r = r - dt2hy * sR * (V(I, J + 1, p) - V(I, J, p));
                }
                const double sedgel = l + dt2 * B.fo(I, J - 1, p, comp);
                const double sedger = r + dt2 * B.fo(I, J, p, comp);
                double res = upwind(sedgel, sedger, V(I, J, p), g.rel_eps);
                res = bc_edge(res, J, g.domlo[1], g.domhi[1], Cb.bclo[1], Cb.bchi[1], vn1, sL, sR,
                              sedgel, sedger);
                B.se[1](I, J, p, comp) = res;
            }
        }
        // z-face q
        if ((q >= z0 && q <= z1) || (q == z1 + 1 && last_z)) {
            if (col && a >= 1 && a <= nx && b >= 1 && b <= ny) {
                double l = slz, r = srz;
                l = edge_side<CONS>(l, dt2hx, dt4hx, U(I + 1, J, p), U(I, J, p), TXY[cp][tid + 1],
                                    TXY[cp][tid]);
                r = edge_side<CONS>(r, dt2hx, dt4hx, U(I + 1, J, q), U(I, J, q), TXY[cq][tid + 1],
                                    TXY[cq][tid]);
                l = edge_side<CONS>(l, dt2hy, dt4hy, V(I, J + 1, p), V(I, J, p), TYX[cp][tid + W],
                                    TYX[cp][tid]);
                r = edge_side<CONS>(r, dt2hy, dt4hy, V(I, J + 1, q), V(I, J, q), TYX[cq][tid + W],
                                    TYX[cq][tid]);
                if (CONS) {
                    l = l - dt2hz * v[3] * (Wm(I, J, q) - Wm(I, J, p));
                    r = r - dt2hz * v[4] * (Wm(I, J, q + 1) - Wm(I, J, q));
                }
                const double sedgel = l + dt2 * B.fo(I, J, p, comp);
                const double sedger = r + dt2 * B.fo(I, J, q, comp);
                double res = upwind(sedgel, sedger, Wm(I, J, q), g.rel_eps);
                res = bc_edge(res, q, g.domlo[2], g.domhi[2], Cb.bclo[2], Cb.bchi[2], vn2, v[3],
                              v[4], sedgel, sedger);
                B.se[2](I, J, q, comp) = res;
            }
        }
        slx_p = slx;
        srx_p = srx;
        sly_p = sly;
        sry_p = sry;
    }
}

}  // namespace

namespace mxh {

bool fused_edge_scal_ok(const mx_adv_ctx* c, int is_conservative) {
    (void)is_conservative;
    return c->p.dim == 3 && (c->p.ppm_type == 1 || c->p.ppm_type == 2) &&
           c->p.ppm_trace_forces == 0 && c->fused_edge;
}

int make_edge_scal_fused(mx_adv_ctx* c, int lev, const mx_multifab* state,
                         const mx_face_multifab* sedge, const mx_face_multifab* umac,
                         const mx_multifab* force, int is_vel, const mx_bcrec* bcs,
                         int start_scomp, int start_bccomp, int num_comp, int is_conservative) {
    constexpr int TX = 32, TY = 8;
    const auto& grids = c->grids[lev];
    const int tz = c->fused_tz > 0 ? c->fused_tz : 32;
    std::vector<FBox> boxes;
    long long maxtiles = 0;
    for (size_t b = 0; b < grids.size(); ++b) {
        FBox B;
        HBox g = hbox(grids[b]);
        for (int q = 0; q < 3; ++q) {
            B.lo[q] = g.lo[q];
            B.hi[q] = g.hi[q];
        }
        B.s = to_da(state->fab[b]);
        B.fo = to_da(force->fab[b]);
        for (int d = 0; d < 3; ++d) {
            B.um[d] = to_da(umac->dir[d].fab[b]);
            B.se[d] = to_da(sedge->dir[d].fab[b]);
        }
        const long long nt = (long long)((g.hi[0] - g.lo[0] + TX) / TX) *
                             ((g.hi[1] - g.lo[1] + TY) / TY) * ((g.hi[2] - g.lo[2] + tz) / tz);
        maxtiles = std::max(maxtiles, nt);
        boxes.push_back(B);
    }
    std::vector<FComp> comps(num_comp);
    for (int vc = 0; vc < num_comp; ++vc)
        for (int d = 0; d < 3; ++d) {
            comps[vc].bclo[d] = bcs[start_bccomp + vc].lo[d];
            comps[vc].bchi[d] = bcs[start_bccomp + vc].hi[d];
        }
    FArgs A;
    A.start_scomp = start_scomp;
    A.num_comp = num_comp;
    A.is_vel = is_vel;
    A.is_conservative = is_conservative;
    A.tz = tz;
    Table tab(c);
    size_t ob = tab.add(boxes);
    size_t oc = tab.add(comps);
    int rc = tab.upload();
    if (rc) return rc;
    const FBox* db = tab.dev<FBox>(ob);
    const FComp* dc = tab.dev<FComp>(oc);
    KGeom kg = kgeom(c, lev);
    dim3 grid((unsigned)maxtiles, (unsigned)grids.size(), (unsigned)num_comp);
    dim3 block(384);
    const bool cons = is_conservative != 0;
    if (c->p.ppm_type == 1) {
        if (cons)
            MX_LAUNCH(c, "k_es_fused", (k_es_fused<1, TX, TY, true>), grid, block, db, dc, kg, A);
        else
            MX_LAUNCH(c, "k_es_fused", (k_es_fused<1, TX, TY, false>), grid, block, db, dc, kg, A);
    } else {
        if (cons)
            MX_LAUNCH(c, "k_es_fused", (k_es_fused<2, TX, TY, true>), grid, block, db, dc, kg, A);
        else
            MX_LAUNCH(c, "k_es_fused", (k_es_fused<2, TX, TY, false>), grid, block, db, dc, kg, A);
    }
    return MX_OK;
}

}  // namespace mxh
