// mx_adjoint.hip -- a differentiable scalar advection step and its
// hand-written adjoint (vector-Jacobian product).
//
// The step, on one fully periodic box [0, N)^3 with a fixed MAC velocity,
// for every scalar component:
//   sedge = MakeEdgeScal(s, umac, force = 0, advective form)   (ppm_type 1)
//   F     = umac * sedge              (MakeRhoXFlux, predict_rhoX,
//                                      MaestroMakeFlux.cpp:118-140)
//   snew  = s + dt * (-div F + 0)     (UpdateScal, MaestroAdvection.cpp:165-183)
// The forward runs the production kernels (periodic ghost fill, the fused
// edge-state kernel) plus an update kernel.
//
// The adjoint differentiates the same arithmetic with every limiter and
// upwind branch frozen at its forward value (the function is piecewise
// smooth; at a switch point the derivative of the branch taken is used).
// It runs as a tape of the forward intermediates (traced states, upwind
// states, corner-coupled states, edge states, on dense periodic N^3 arrays)
// and reverse sweeps that mirror the forward stages.  Every reverse stage
// gathers the adjoints of its inputs from the adjoints of the outputs that
// use them, so no atomics are used and the result is deterministic.
// Adjoints that are cheap to recompute from their neighbours' inputs are
// recomputed where they are gathered instead of stored:
//   B1 flux, update and edge-state upwinding -> per-cell edge-term adjoints
//      (in the edge-state tape kernel: it needs only gbar and the velocity)
//   B4 corner terms (recomputed)    -> per-cell transverse-term adjoints
//   B3 + B5 the face adjoints of the predictor states and velocity: flux,
//      corner and transverse terms, summed where B6 and B10 read them
//   B6 characteristic trace and CW limiter (per cell and direction)
//   B7 PPM face values          B8 van Leer slopes   B9 state adjoint
//   B10 velocity adjoint summed over components.
#include <algorithm>
#include <cstring>

#include "mx_internal.H"

using namespace mxd;
using namespace mxh;

namespace {

constexpr int NTB = 256;

constexpr double C23 = 2.0 / 3.0;

__host__ __device__ constexpr int pidx(int d, int t) { return d * 2 + (t > d ? t - 1 : t); }

// periodic indexing of a dense N^3 array (N^3 fits 31 bits; checked on the
// host, so all index arithmetic is 32-bit).  P2: N is a power of two and a
// coordinate within one period of [0, N) wraps with a mask.
template <bool P2>
struct PerT {
    int n, m, lg;
    __device__ __forceinline__ int w(int x) const {
        if (P2) return x & m;
        return x < 0 ? x + n : (x >= n ? x - n : x);
    }
    __device__ __forceinline__ int at(int i, int j, int k) const {
        if (P2) return w(i) | (w(j) << lg) | (w(k) << (2 * lg));
        return w(i) + n * (w(j) + n * w(k));
    }
    // neighbour of (i,j,k) at offset o along direction d
    __device__ __forceinline__ int nb(int i, int j, int k, int d, int o) const {
        return at(i + (d == 0 ? o : 0), j + (d == 1 ? o : 0), k + (d == 2 ? o : 0));
    }
};

template <bool P2>
__device__ __forceinline__ PerT<P2> make_per(int n) {
    PerT<P2> P;
    P.n = n;
    P.m = n - 1;
    P.lg = 31 - __clz(n);
    return P;
}

struct AB {
    const double* S;   // state, ncomp x N^3
    const double* U[3];
    const double* G;   // adjoint of snew
    double *IP[3], *IM[3], *SIMH[3], *COR[6], *SE[3];
    double *EB[6], *TB[3];
    double *SLB[3], *SRB[3];  // face sums (k_adj_fsum), when used
    double* UBC[3];           // per-component face velocity adjoints (k_adj_trace)
    double *S0S, *FLB[3], *FRB[3], *UHI[3], *ULO[3];  // S0S: trace terms of s, summed over d
    double* DV[3][4];  // ppm_type 2: limiter adjoints of s(p - 2d, p - d, p + d, p + 2d)
    double *DSA[3], *DSB[3], *DLA[3], *DLB[3], *QM[3], *Q0[3], *QP[3];
    DA SBF;            // state adjoint: the caller's FAB
    long long cs;      // N^3
    int n, ncomp;
    double dt, rel_eps, dx[3];
    // per-direction factors, evaluated once on the host by the same
    // expressions: dt/h, (dt/4)/h (edge terms), (dt/6)/h (transverse terms)
    double fdt[3], f4[3], f6[3];
};

__device__ __forceinline__ bool point(const AB& A, int& i, int& j, int& k, long long& p) {
    const unsigned q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (unsigned)A.cs) return false;
    const unsigned n = (unsigned)A.n;
    const unsigned r = q / n;
    i = (int)(q - r * n);
    k = (int)(r / n);
    j = (int)(r - (unsigned)k * n);
    p = q;
    return true;
}

__device__ __forceinline__ void weights(double u, double thr, double& wl, double& wr) {
    if (!(dabs(u) > thr)) {
        wl = 0.5;
        wr = 0.5;
    } else if (u > 0.0) {
        wl = 1.0;
        wr = 0.0;
    } else {
        wl = 0.0;
        wr = 1.0;
    }
}

__device__ __forceinline__ double upwind(double l, double r, double u, double thr) {
    double res = (u > 0.0) ? l : r;
    res = (dabs(u) > thr) ? res : 0.5 * (l + r);
    return res;
}

__device__ __forceinline__ double dabs_d(double x) { return signbit(x) ? -1.0 : 1.0; }

// ---------------------------------------------------------------- forward tape
// A1 + A2: traced PPM states of every cell and direction, and the upwind
// state simh on the face below the cell (from the traced state of the cell
// below it, recomputed here)
template <int PPM, int D, class PT>
__device__ __forceinline__ void ppm_cell(const AB& A, const PT& P, const double* S, int i, int j,
                                         int k, int pos, double& Ip, double& Im) {
    constexpr int G = PPM == 2 ? 3 : 2;
    double v[9];
#pragma unroll
    for (int o = -4; o <= 4; ++o) v[4 + o] = (o >= -G && o <= G) ? S[P.nb(i, j, k, D, o)] : 0.0;
    double sm, sp;
    ppm_edges<PPM, D>(v, pos, 0, A.n - 1, MX_BC_INT_DIR, MX_BC_INT_DIR, sm, sp);
    ppm_trace(v[4], sm, sp, A.U[D][P.nb(i, j, k, D, 1)], A.U[D][P.at(i, j, k)], A.dt, A.dx[D],
              A.rel_eps, Ip, Im);
}

template <int PPM, int D, class PT>
__device__ __forceinline__ void tape_dir(const AB& A, const PT& P, const double* S, int i, int j,
                                         int k, long long p, double& ip, double& im, double& h) {
    const int pos = D == 0 ? i : (D == 1 ? j : k);
    ppm_cell<PPM, D>(A, P, S, i, j, k, pos, ip, im);
    double ipm, imm;
    ppm_cell<PPM, D>(A, P, S, i - (D == 0), j - (D == 1), k - (D == 2), pos - 1, ipm, imm);
    h = upwind(ipm, im, A.U[D][p], 0.0);
}

template <int PPM, bool P2>
__global__ void __launch_bounds__(NTB) k_tape_ppm(AB A) {
    int i, j, k;
    long long p;
    if (!point(A, i, j, k, p)) return;
    const PerT<P2> P = make_per<P2>(A.n);
    const long long co = (long long)blockIdx.y * A.cs;
    const double* S = A.S + co;
    double ip[3], imv[3], h[3];
    tape_dir<PPM, 0>(A, P, S, i, j, k, p, ip[0], imv[0], h[0]);
    tape_dir<PPM, 1>(A, P, S, i, j, k, p, ip[1], imv[1], h[1]);
    tape_dir<PPM, 2>(A, P, S, i, j, k, p, ip[2], imv[2], h[2]);
    // every result is stored after the last load (stores and loads share
    // one in-order counter, so a load issued after a store waits for it)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        A.IP[d][co + p] = ip[d];
        A.IM[d][co + p] = imv[d];
        A.SIMH[d][co + p] = h[d];
    }
}

// transverse term of cell c in direction t (dt/6h_t (u+ + u-) (simh+ - simh-))
template <class PT>
__device__ __forceinline__ double tt(const AB& A, const PT& P, long long co, int i, int j, int k,
                                     int t) {
    const long long c = P.at(i, j, k);
    const long long c1 = P.nb(i, j, k, t, 1);
    return A.f6[t] * (A.U[t][c1] + A.U[t][c]) * (A.SIMH[t][co + c1] - A.SIMH[t][co + c]);
}

// A3: corner-coupled states on d-faces, corrected by t
template <bool P2>
__global__ void __launch_bounds__(NTB) k_tape_corner(AB A) {
    int i, j, k;
    long long p;
    if (!point(A, i, j, k, p)) return;
    const PerT<P2> P = make_per<P2>(A.n);
    const long long co = (long long)blockIdx.y * A.cs;
    double cor[6];
    for (int d = 0; d < 3; ++d) {
        const int im = i - (d == 0), jm = j - (d == 1), km = k - (d == 2);
        for (int t = 0; t < 3; ++t) {
            if (t == d) continue;
            double l = A.IP[d][co + P.at(im, jm, km)];
            double r = A.IM[d][co + p];
            l = l - tt(A, P, co, im, jm, km, t);
            r = r - tt(A, P, co, i, j, k, t);
            cor[pidx(d, t)] = upwind(l, r, A.U[d][p], A.rel_eps);
        }
    }
    for (int q = 0; q < 6; ++q) A.COR[q][co + p] = cor[q];
}

// edge term of cell c for the corner state on t-faces corrected by o
template <class PT>
__device__ __forceinline__ double et(const AB& A, const PT& P, long long co, int i, int j, int k,
                                     int t, int o) {
    const long long c = P.at(i, j, k);
    const long long c1 = P.nb(i, j, k, t, 1);
    const double* K = A.COR[pidx(t, o)] + co;
    return A.f4[t] * (A.U[t][c1] + A.U[t][c]) * (K[c1] - K[c]);
}

// adjoint of the edge state on d-face c through the flux (Fb u) and its
// upwinding weights
template <class PT>
__device__ __forceinline__ double edge_bar(const AB& A, const PT& P, long long co, int i, int j,
                                           int k, int d, double& wl, double& wr) {
    const long long c = P.at(i, j, k);
    const double Fb = A.fdt[d] * (A.G[co + c] - A.G[co + P.nb(i, j, k, d, -1)]);
    const double u = A.U[d][c];
    weights(u, A.rel_eps, wl, wr);
    return Fb * u;
}

// A4: edge states
template <bool P2>
__global__ void __launch_bounds__(NTB) k_tape_edge(AB A) {
    int i, j, k;
    long long p;
    if (!point(A, i, j, k, p)) return;
    const PerT<P2> P = make_per<P2>(A.n);
    const long long co = (long long)blockIdx.y * A.cs;
    const double dt2 = 0.5 * A.dt;
    double se[3];
    for (int d = 0; d < 3; ++d) {
        const int im = i - (d == 0), jm = j - (d == 1), km = k - (d == 2);
        double l = A.IP[d][co + P.at(im, jm, km)];
        double r = A.IM[d][co + p];
        for (int t = 0; t < 3; ++t) {
            if (t == d) continue;
            const int o = 3 - d - t;
            l = l - et(A, P, co, im, jm, km, t, o);
            r = r - et(A, P, co, i, j, k, t, o);
        }
        const double sedgel = l + dt2 * 0.0;
        const double sedger = r + dt2 * 0.0;
        se[d] = upwind(sedgel, sedger, A.U[d][p], A.rel_eps);
    }
    // B1 (it needs only the output adjoint and the velocity): the flux,
    // update and edge-state upwinding give the per-cell edge-term adjoints;
    // each edge term of cell p enters the left state of d-face p + d and the
    // right state of d-face p, d = 3-t-o
    double lb1[3], rb0[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double wl, wr;
        const double seb = edge_bar(A, P, co, i, j, k, d, wl, wr);
        rb0[d] = wr * seb;
        double wl1, wr1;
        const double seb1 = edge_bar(A, P, co, i + (d == 0), j + (d == 1), k + (d == 2), d, wl1, wr1);
        lb1[d] = wl1 * seb1;
    }
    for (int d = 0; d < 3; ++d) A.SE[d][co + p] = se[d];
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
        for (int o = 0; o < 3; ++o) {
            if (o == t) continue;
            const int d = 3 - t - o;
            A.EB[pidx(t, o)][co + p] = -lb1[d] - rb0[d];
        }
}

// ---------------------------------------------------------------- reverse
// B3 + B5: the adjoints on t-face (i,j,k) of its left and right predictor
// states (slb, srb) and, with want_u, of its velocity for this component
// (ub): each the sum, in this order, of the flux term (edge-state
// upwinding), the corner-state terms (edge terms, corner upwinding) and the
// transverse term (predictor upwinding)
template <class PT>
__device__ __forceinline__ void face_totals(const AB& A, const PT& P, long long co, int i, int j,
                                            int k, int t, bool want_u, double& slb, double& srb,
                                            double& ub) {
    const long long f = P.at(i, j, k);
    const long long fm = P.nb(i, j, k, t, -1), fp = P.nb(i, j, k, t, 1);
    const double u0 = A.U[t][f], um = A.U[t][fm], up = A.U[t][fp];
    // flux
    const double Fb = A.fdt[t] * (A.G[co + f] - A.G[co + fm]);
    const double seb = Fb * u0;
    double wl, wr;
    weights(u0, A.rel_eps, wl, wr);
    ub = want_u ? Fb * A.SE[t][co + f] : 0.0;
    slb = wl * seb;
    srb = wr * seb;
    // corner states on t-faces, through the edge terms of the cells below
    // and above the face
    const double b = A.f4[t];
    double ubc = 0.0, slbc = 0.0, srbc = 0.0;
#pragma unroll
    for (int o = 0; o < 3; ++o) {
        if (o == t) continue;
        const int q = pidx(t, o);
        const double ebm = A.EB[q][co + fm];
        const double eb0 = A.EB[q][co + f];
        const double kb = ebm * b * (u0 + um) - eb0 * b * (up + u0);
        if (want_u) {
            const double* K = A.COR[q] + co;
            ubc += ebm * b * (K[f] - K[fm]) + eb0 * b * (K[fp] - K[f]);
        }
        slbc += wl * kb;
        srbc += wr * kb;
    }
    ub += ubc;
    slb += slbc;
    srb += srbc;
    // upwind state simh, through the transverse terms of the cells below and
    // above the face
    const double a = A.f6[t];
    const double tbm = A.TB[t][co + fm], tb0 = A.TB[t][co + f];
    const double hb = tbm * a * (u0 + um) - tb0 * a * (up + u0);
    if (want_u) {
        const double* H = A.SIMH[t] + co;
        ub = ub + (tbm * a * (H[f] - H[fm]) + tb0 * a * (H[fp] - H[f]));
    }
    double wl0, wr0;
    weights(u0, 0.0, wl0, wr0);
    slb = slb + wl0 * hb;
    srb = srb + wr0 * hb;
}

// adjoint of the corner state q on d-face (i,j,k) (through the edge terms of
// the cells on both sides of it), with its upwinding weights: the left and
// right predictor-state adjoints it passes on
template <class PT>
__device__ __forceinline__ void corner_bar(const AB& A, const PT& P, long long co, int i, int j,
                                           int k, int d, int q, double& lb, double& rb) {
    const double b = A.f4[d];
    const long long c = P.at(i, j, k);
    const long long fm = P.nb(i, j, k, d, -1), fp = P.nb(i, j, k, d, 1);
    const double u0 = A.U[d][c], um = A.U[d][fm], up = A.U[d][fp];
    const double ebm = A.EB[q][co + fm];
    const double eb0 = A.EB[q][co + c];
    const double kb = ebm * b * (u0 + um) - eb0 * b * (up + u0);
    double wl, wr;
    weights(u0, A.rel_eps, wl, wr);
    lb = wl * kb;
    rb = wr * kb;
}

// B4: per-cell transverse-term adjoints: the transverse term of cell p in
// direction o enters the corner states on the d-faces p (right state) and
// p + d (left state), d != o
template <bool P2>
__global__ void __launch_bounds__(NTB) k_adj_tterm(AB A) {
    int i, j, k;
    long long p;
    if (!point(A, i, j, k, p)) return;
    const PerT<P2> P = make_per<P2>(A.n);
    const long long co = (long long)blockIdx.y * A.cs;
    double tb[3];
    for (int o = 0; o < 3; ++o) {
        double s = 0.0;
        for (int d = 0; d < 3; ++d) {
            if (d == o) continue;
            const int q = pidx(d, o);
            double lb1, rb1, lb0, rb0;
            corner_bar(A, P, co, i + (d == 0), j + (d == 1), k + (d == 2), d, q, lb1, rb1);
            corner_bar(A, P, co, i, j, k, d, q, lb0, rb0);
            s += lb1 + rb0;
        }
        tb[o] = -s;
    }
    for (int o = 0; o < 3; ++o) A.TB[o][co + p] = tb[o];
}

// B3 + B5 stored once per face: the left and right predictor-state adjoints
// of every d-face (for the trace variant that reads them)
template <bool P2>
__global__ void __launch_bounds__(NTB) k_adj_fsum(AB A) {
    int i, j, k;
    long long p;
    if (!point(A, i, j, k, p)) return;
    const PerT<P2> P = make_per<P2>(A.n);
    const long long co = (long long)blockIdx.y * A.cs;
    double l[3], r[3], unused;
#pragma unroll
    for (int d = 0; d < 3; ++d) face_totals(A, P, co, i, j, k, d, false, l[d], r[d], unused);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        A.SLB[d][co + p] = l[d];
        A.SRB[d][co + p] = r[d];
    }
}

// adjoint of the characteristic trace (ppm_trace) of a cell: from the
// adjoints Ipb, Imb of its traced states to its sm, sp, s0 and the velocities
// of its upper and lower faces
struct TraceBar {
    double smb, spb, s0b, uhib, ulob;
};
__device__ __forceinline__ TraceBar trace_bar(const AB& A, double s0, double sm, double sp,
                                              double s6, double uhi, double ulo, double h,
                                              double Ipb, double Imb) {
    TraceBar r = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (uhi > A.rel_eps) {
        const double sg = dabs(uhi) * A.dt / h;
        const double f = 1.0 - C23 * sg;
        const double Aq = sp - sm - f * s6;
        r.spb += Ipb * (1.0 - 0.5 * sg * (1.0 + 3.0 * f));
        r.smb += Ipb * (-0.5 * sg * (-1.0 + 3.0 * f));
        r.s0b += Ipb * (3.0 * sg * f);
        const double sgb = Ipb * (-0.5 * Aq - 0.5 * sg * (C23 * s6));
        r.uhib += sgb * dabs_d(uhi) * A.dt / h;
    } else {
        r.s0b += Ipb;
    }
    if (ulo < -A.rel_eps) {
        const double sg = dabs(ulo) * A.dt / h;
        const double f = 1.0 - C23 * sg;
        const double Bq = sp - sm + f * s6;
        r.smb += Imb * (1.0 + 0.5 * sg * (-1.0 - 3.0 * f));
        r.spb += Imb * (0.5 * sg * (1.0 - 3.0 * f));
        r.s0b += Imb * (3.0 * sg * f);
        const double sgb = Imb * (0.5 * Bq - 0.5 * sg * (C23 * s6));
        r.ulob += sgb * dabs_d(ulo) * A.dt / h;
    } else {
        r.s0b += Imb;
    }
    return r;
}

// B6: characteristic trace and CW limiter of cell p in direction d (FS: the
// face adjoints are read from k_adj_fsum's arrays instead of recomputed)
template <bool P2, bool FS, bool ROLL>
__global__ void __launch_bounds__(NTB) k_adj_trace(AB A) {
    int i, j, k;
    long long p;
    if (!point(A, i, j, k, p)) return;
    const PerT<P2> P = make_per<P2>(A.n);
    const long long co = (long long)blockIdx.y * A.cs;
    const double* S = A.S + co;
    // ROLL: one direction per loop trip with its stores at the end of the
    // trip (fewer registers, more waves); else all loads before any store
    double r0[3], r1[3], r2[3], r3[3], r4[3], ubd[3] = {0.0, 0.0, 0.0};
    double s0s = 0.0;  // the trace's direct terms of s, summed over d in order
#pragma unroll(ROLL ? 1 : 3)
    for (int d = 0; d < 3; ++d) {
        double v[9];
        for (int o = -4; o <= 4; ++o) v[4 + o] = (o >= -2 && o <= 2) ? S[P.nb(i, j, k, d, o)] : 0.0;
        // forward, keeping the face values before the limiter
        const double dl = vl_slope(v[2], v[3], v[4]);
        const double dc = vl_slope(v[3], v[4], v[5]);
        const double dr = vl_slope(v[4], v[5], v[6]);
        double sm0 = 0.5 * (v[4] + v[3]) - (dc - dl) / 6.0;
        sm0 = dmax(sm0, dmin(v[4], v[3]));
        sm0 = dmin(sm0, dmax(v[4], v[3]));
        double sp0 = 0.5 * (v[5] + v[4]) - (dr - dc) / 6.0;
        sp0 = dmax(sp0, dmin(v[5], v[4]));
        sp0 = dmin(sp0, dmax(v[5], v[4]));
        const double s0 = v[4];
        int br;
        double sm = sm0, sp = sp0;
        if ((sp - s0) * (s0 - sm) <= 0.0) {
            br = 0;
            sp = s0;
            sm = s0;
        } else if (dabs(sp - s0) >= 2.0 * dabs(sm - s0)) {
            br = 1;
            sp = 3.0 * s0 - 2.0 * sm;
        } else if (dabs(sm - s0) >= 2.0 * dabs(sp - s0)) {
            br = 2;
            sm = 3.0 * s0 - 2.0 * sp;
        } else {
            br = 3;
        }
        const long long fhi = P.nb(i, j, k, d, 1);
        const double uhi = A.U[d][fhi], ulo = A.U[d][p];
        const double h = A.dx[d];
        const double s6 = 6.0 * s0 - 3.0 * (sm + sp);
        double Ipb, Imb;
        if (FS) {
            Ipb = A.SLB[d][co + fhi];
            Imb = A.SRB[d][co + p];
        } else {
            double unused0, unused1, unused2;
            face_totals(A, P, co, i + (d == 0), j + (d == 1), k + (d == 2), d, false, Ipb,
                        unused0, unused1);
            // face p's totals also give its velocity adjoint for this component
            face_totals(A, P, co, i, j, k, d, true, unused2, Imb, ubd[ROLL ? 0 : d]);
        }
        const TraceBar tr = trace_bar(A, s0, sm, sp, s6, uhi, ulo, h, Ipb, Imb);
        const double smb = tr.smb, spb = tr.spb, uhib = tr.uhib, ulob = tr.ulob;
        double s0b = tr.s0b;
        // CW limiter
        double sm0b = 0.0, sp0b = 0.0;
        if (br == 0) {
            s0b += smb + spb;
        } else if (br == 1) {
            s0b += 3.0 * spb;
            sm0b = smb - 2.0 * spb;
        } else if (br == 2) {
            s0b += 3.0 * smb;
            sp0b = spb - 2.0 * smb;
        } else {
            sm0b = smb;
            sp0b = spb;
        }
        if (ROLL) {
            s0s = d == 0 ? s0b : s0s + s0b;  // (t0 + t1) + t2, as unrolled
            A.FLB[d][co + p] = sm0b;
            A.FRB[d][co + p] = sp0b;
            A.UHI[d][co + p] = uhib;
            if (FS)
                A.ULO[d][co + p] = ulob;
            else
                A.UBC[d][co + p] = ubd[0] + ulob;  // both belong to face p
            continue;
        }
        r0[d] = s0b;
        r1[d] = sm0b;
        r2[d] = sp0b;
        r3[d] = uhib;
        r4[d] = ulob;
    }
    if (ROLL) {
        A.S0S[co + p] = s0s;
        return;
    }
    A.S0S[co + p] = (r0[0] + r0[1]) + r0[2];
    for (int d = 0; d < 3; ++d) {
        A.FLB[d][co + p] = r1[d];  // face p (below the cell)
        A.FRB[d][co + p] = r2[d];  // face above the cell
        A.UHI[d][co + p] = r3[d];
        // the face-p velocity adjoints of this component: through the face
        // totals and through this cell's trace (its lower face)
        if (FS)
            A.ULO[d][co + p] = r4[d];
        else
            A.UBC[d][co + p] = ubd[d] + r4[d];
    }
}

// ---------------------------------------------------------------- ppm_type 2
// gradient of the Colella-Sekora face value cs_face(sa, sb, sc, sd) with
// respect to its four cells, on the branch the forward takes
__device__ __forceinline__ void cs_face_grad(double sa, double sb, double sc, double sd,
                                             double g[4]) {
    const double C = 1.25;
    const double e = (7.0 / 12.0) * (sb + sc) - (1.0 / 12.0) * (sa + sd);
    g[0] = -1.0 / 12.0;
    g[1] = 7.0 / 12.0;
    g[2] = 7.0 / 12.0;
    g[3] = -1.0 / 12.0;
    if ((e - sb) * (sc - e) < 0.0) {
        // e = (sb + sc) / 2 - D2LIM / 6 with D2LIM = sgn * max(min(C sgn D2L,
        // min(C sgn D2R, sgn D2)), 0); sgn is piecewise constant
        const double D2 = 3.0 * (sb - 2.0 * e + sc);
        const double D2L = sa - 2.0 * sb + sc;
        const double D2R = sb - 2.0 * sc + sd;
        const double sgn = dsign1(D2);
        const double a1 = C * sgn * D2L, a2 = C * sgn * D2R, a3 = sgn * D2;
        const double inner = dmin(a2, a3);
        const bool in3 = a3 < a2;          // dmin(a2, a3) returns a3
        const double outer = dmin(a1, inner);
        const bool out_in = inner < a1;    // dmin(a1, inner) returns inner
        const bool zero = outer < 0.0;     // dmax(outer, 0) returns 0
        // d D2LIM / d(sa, sb, sc, sd); sgn^2 = 1
        double l[4] = {0.0, 0.0, 0.0, 0.0};
        if (!zero) {
            if (!out_in) {
                l[0] = C; l[1] = -2.0 * C; l[2] = C;                       // C D2L
            } else if (!in3) {
                l[1] = C; l[2] = -2.0 * C; l[3] = C;                       // C D2R
            } else {
                l[0] = 0.5; l[1] = -0.5; l[2] = -0.5; l[3] = 0.5;          // D2
            }
        }
        g[0] = -l[0] / 6.0;
        g[1] = 0.5 - l[1] / 6.0;
        g[2] = 0.5 - l[2] / 6.0;
        g[3] = -l[3] / 6.0;
    }
}

// adjoint of the Colella-Sekora cell limiter (interior: threshold 1e-10) of
// the cell with v[-2..2] = s(cell-2 .. cell+2) and faces fl, f0, fr, frr
// (cell-1 .. cell+2): from smb, spb to the cells (vb[-2..2]) and to the
// faces f0, fr (fl and frr enter only branch decisions)
__device__ __forceinline__ void cs_cell_bar(const double* v, double fl, double f0, double fr,
                                            double frr, double smb, double spb, double* vb,
                                            double& f0b, double& frb) {
    const double C = 1.25;
    const double s0 = v[0];
    const double ap0 = fr - s0;
    const double am0 = f0 - s0;
    const bool bigp = dabs(ap0) > 2.0 * dabs(am0);
    const bool bigm = dabs(am0) > 2.0 * dabs(ap0);
    bool extremum = false;
    if (am0 * ap0 >= 0.0) {
        extremum = true;
    } else if (bigp || bigm) {
        const double dafacem = f0 - fl, dafacep = frr - fr;
        const double dabarm = s0 - v[-1], dabarp = v[1] - s0;
        const double dafacemin = dmin(dabs(dafacem), dabs(dafacep));
        const double dabarmin = dmin(dabs(dabarm), dabs(dabarp));
        const bool face = dafacemin >= dabarmin;
        const double dachkm = face ? dafacem : dabarm;
        const double dachkp = face ? dafacep : dabarp;
        extremum = dachkm * dachkp <= 0.0;
    }
    for (int o = -2; o <= 2; ++o) vb[o] = 0.0;
    // sm = s0 + alpham, sp = s0 + alphap
    vb[0] = smb + spb;
    double amb = 0.0, apb = 0.0;  // adjoints of am0, ap0
    if (extremum) {
        const double D2 = 6.0 * (am0 + ap0);
        const double D2L = v[-2] - 2.0 * v[-1] + s0;
        const double D2R = s0 - 2.0 * v[1] + v[2];
        const double D2C = v[-1] - 2.0 * s0 + v[1];
        const double sgn = dsign1(D2);
        const double t0 = sgn * D2, t1 = C * sgn * D2L, t2 = C * sgn * D2R, t3 = C * sgn * D2C;
        // D2LIM = dmax(dmin(t0, dmin(t1, dmin(t2, t3))), 0)
        const double m23 = dmin(t2, t3);
        const int i23 = t3 < t2 ? 3 : 2;
        const double m123 = dmin(t1, m23);
        const int i123 = m23 < t1 ? i23 : 1;
        const double m0123 = dmin(t0, m123);
        const int imin = m123 < t0 ? i123 : 0;
        const bool lzero = m0123 < 0.0;
        const double D2LIM = lzero ? 0.0 : m0123;
        const double D2ABS = dmax(dabs(D2), 1.e-10);
        const bool absd = !(dabs(D2) < 1.e-10);  // dmax returns |D2|
        // alpham = am0 * D2LIM / D2ABS, alphap = ap0 * D2LIM / D2ABS
        amb = smb * D2LIM / D2ABS;
        apb = spb * D2LIM / D2ABS;
        const double wsum = smb * am0 + spb * ap0;
        const double limb = wsum / D2ABS;
        const double absb = -wsum * D2LIM / (D2ABS * D2ABS);
        double D2b = absd ? absb * sgn : 0.0;  // d|D2|/dD2 = sgn
        if (!lzero) {
            if (imin == 0) {
                D2b += limb * sgn;
            } else {
                const double w = limb * C * sgn;
                if (imin == 1) {
                    vb[-2] += w;
                    vb[-1] -= 2.0 * w;
                    vb[0] += w;
                } else if (imin == 2) {
                    vb[0] += w;
                    vb[1] -= 2.0 * w;
                    vb[2] += w;
                } else {
                    vb[-1] += w;
                    vb[0] -= 2.0 * w;
                    vb[1] += w;
                }
            }
        }
        amb += 6.0 * D2b;
        apb += 6.0 * D2b;
    } else {
        amb = smb;
        apb = spb;
        if (bigp) {
            const double sgn = dsign1(am0);
            const double amax = -ap0 * ap0 / (4.0 * (am0 + ap0));
            const double delam = v[-1] - s0;
            if (sgn * amax >= sgn * delam) {
                // alphap is replaced
                apb = 0.0;
                if (sgn * (delam - am0) >= 1.e-10) {
                    const double q = delam * delam - delam * am0;
                    const double r = sqrt(q);
                    double delb = -2.0 * spb;
                    if (r > 0.0) {
                        delb += -spb * sgn * (2.0 * delam - am0) / r;
                        amb += spb * sgn * delam / r;
                    }
                    vb[-1] += delb;
                    vb[0] -= delb;
                } else {
                    amb += -2.0 * spb;
                }
            }
        }
        if (bigm) {
            const double sgn = dsign1(ap0);
            const double amax = -am0 * am0 / (4.0 * (am0 + ap0));
            const double delap = v[1] - s0;
            if (sgn * amax >= sgn * delap) {
                amb = 0.0;
                if (sgn * (delap - ap0) >= 1.e-10) {
                    const double q = delap * delap - delap * ap0;
                    const double r = sqrt(q);
                    double delb = -2.0 * smb;
                    if (r > 0.0) {
                        delb += -smb * sgn * (2.0 * delap - ap0) / r;
                        apb += smb * sgn * delap / r;
                    }
                    vb[1] += delb;
                    vb[0] -= delb;
                } else {
                    apb += -2.0 * smb;
                }
            }
        }
    }
    // am0 = f0 - s0, ap0 = fr - s0
    f0b = amb;
    frb = apb;
    vb[0] -= amb + apb;
}

// B6 for ppm_type 2: trace and Colella-Sekora limiter of cell p in every
// direction.  Per direction it stores the adjoints of its faces p (FLB) and
// p + d (FRB), the direct adjoints of the cells p - 2d .. p + 2d other than p
// (DV[d][0..3] for offsets -2, -1, 1, 2), the velocity terms, and adds its
// direct term of s(p) to S0S.
template <bool P2>
__global__ void __launch_bounds__(NTB) k_adj_trace2(AB A) {
    int i, j, k;
    long long p;
    if (!point(A, i, j, k, p)) return;
    const PerT<P2> P = make_per<P2>(A.n);
    const long long co = (long long)blockIdx.y * A.cs;
    const double* S = A.S + co;
    double s0s = 0.0;
#pragma unroll 1
    for (int d = 0; d < 3; ++d) {
        double w[7];
        for (int o = -3; o <= 3; ++o) w[3 + o] = S[P.nb(i, j, k, d, o)];
        const double fl = cs_face(w[0], w[1], w[2], w[3]);
        const double f0 = cs_face(w[1], w[2], w[3], w[4]);
        const double fr = cs_face(w[2], w[3], w[4], w[5]);
        const double frr = cs_face(w[3], w[4], w[5], w[6]);
        double v9[9];
        for (int o = -4; o <= 4; ++o) v9[4 + o] = (o >= -3 && o <= 3) ? w[3 + o] : 0.0;
        double sm, sp;
        cs_cell(v9, fl, f0, fr, frr, 1.e-10, sm, sp);
        const double s0 = w[3];
        const long long fhi = P.nb(i, j, k, d, 1);
        const double uhi = A.U[d][fhi], ulo = A.U[d][p];
        const double s6 = 6.0 * s0 - 3.0 * (sm + sp);
        double Ipb, Imb, ubd, unused0, unused1, unused2;
        face_totals(A, P, co, i + (d == 0), j + (d == 1), k + (d == 2), d, false, Ipb, unused0,
                    unused1);
        face_totals(A, P, co, i, j, k, d, true, unused2, Imb, ubd);
        const TraceBar tr = trace_bar(A, s0, sm, sp, s6, uhi, ulo, A.dx[d], Ipb, Imb);
        const double smb = tr.smb, spb = tr.spb, s0b = tr.s0b, uhib = tr.uhib, ulob = tr.ulob;
        double vb5[5], f0b, frb;
        cs_cell_bar(w + 3, fl, f0, fr, frr, smb, spb, vb5 + 2, f0b, frb);
        s0s = d == 0 ? s0b + vb5[2] : s0s + (s0b + vb5[2]);
        A.FLB[d][co + p] = f0b;
        A.FRB[d][co + p] = frb;
        A.DV[d][0][co + p] = vb5[0];
        A.DV[d][1][co + p] = vb5[1];
        A.DV[d][2][co + p] = vb5[3];
        A.DV[d][3][co + p] = vb5[4];
        A.UHI[d][co + p] = uhib;
        A.UBC[d][co + p] = ubd + ulob;
    }
    A.S0S[co + p] = s0s;
}

// B7-B9 for ppm_type 2: the state adjoint of cell p gathers, per direction,
// the direct limiter terms of the cells p - 2d .. p + 2d and the face
// adjoints of the faces p - d .. p + 2d through their Colella-Sekora values
template <bool P2>
__global__ void __launch_bounds__(NTB) k_adj_state2(AB A) {
    int i, j, k;
    long long p;
    if (!point(A, i, j, k, p)) return;
    const PerT<P2> P = make_per<P2>(A.n);
    const long long co = (long long)blockIdx.y * A.cs;
    const double* S = A.S + co;
    double s = A.G[co + p] + A.S0S[co + p];
#pragma unroll 1
    for (int d = 0; d < 3; ++d) {
        // direct limiter terms: cell c gives DV[o](c) to s(c + o)
        s += A.DV[d][0][co + P.nb(i, j, k, d, 2)] + A.DV[d][1][co + P.nb(i, j, k, d, 1)] +
             A.DV[d][2][co + P.nb(i, j, k, d, -1)] + A.DV[d][3][co + P.nb(i, j, k, d, -2)];
        // faces f = p - d .. p + 2d (f: cells f - d | f, value from f - 2d .. f + d)
        double w[7];
        for (int o = -3; o <= 3; ++o) w[3 + o] = S[P.nb(i, j, k, d, o)];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int fo = m - 1;  // face offset from p
            const long long f = P.nb(i, j, k, d, fo), fm = P.nb(i, j, k, d, fo - 1);
            const double fb = A.FLB[d][co + f] + A.FRB[d][co + fm];
            double g[4];
            // inputs s(f - 2d .. f + d) = w[3 + fo - 2 .. 3 + fo + 1]
            cs_face_grad(w[1 + fo], w[2 + fo], w[3 + fo], w[4 + fo], g);
            // p is input 2 - fo of face p + fo
            s += fb * g[2 - fo];
        }
    }
    A.SBF(i, j, k, (int)blockIdx.y) = s;
}

// B7: adjoints through the PPM type 1 face value of d-face (i,j,k) (cells
// below | above it): of the two cells' values (bsa, bsb) and of their van
// Leer slopes (bla, blb)
template <class PT>
__device__ __forceinline__ void face_bar(const AB& A, const PT& P, long long co,
                                         const double* S, int i, int j, int k, int d, double& bsa,
                                         double& bsb, double& bla, double& blb) {
    const long long c = P.at(i, j, k);
    const long long cm = P.nb(i, j, k, d, -1);
    const double eb = A.FLB[d][co + c] + A.FRB[d][co + cm];
    const double sa = S[cm], sb = S[c];
    const double sa2 = S[P.nb(i, j, k, d, -2)], sb2 = S[P.nb(i, j, k, d, 1)];
    const double dla = vl_slope(sa2, sa, sb), dlb = vl_slope(sa, sb, sb2);
    const double e0 = 0.5 * (sb + sa) - (dlb - dla) / 6.0;
    const double m1 = dmin(sb, sa);
    const bool m1a = sa < sb;  // dmin(sb, sa) returns sa
    const double e1 = dmax(e0, m1);
    const bool e1m = e0 < m1;
    const double M = dmax(sb, sa);
    const bool Ma = sb < sa;  // dmax(sb, sa) returns sa
    bsa = 0.0;
    bsb = 0.0;
    bla = 0.0;
    blb = 0.0;
    if (M < e1) {
        if (Ma) bsa = eb;
        else bsb = eb;
    } else if (e1m) {
        if (m1a) bsa = eb;
        else bsb = eb;
    } else {
        bsa = 0.5 * eb;
        bsb = 0.5 * eb;
        bla = eb / 6.0;
        blb = -eb / 6.0;
    }
}

// B8: adjoints of the van Leer slope of cell (i,j,k) along d, whose adjoint
// is rb, to its three stencil values
template <class PT>
__device__ __forceinline__ void slope_bar(const PT& P, const double* S, int i, int j, int k,
                                          int d, double rb, double& qm, double& q0, double& qp) {
    const double a = S[P.nb(i, j, k, d, -1)], b = S[P.at(i, j, k)], c = S[P.nb(i, j, k, d, 1)];
    const double dsc = 0.5 * (c - a);
    const double dsl = 2.0 * (b - a);
    const double dsr = 2.0 * (c - b);
    // the slope is sign(dsc) * min(|dsc|, min(|dsl|, |dsr|)) when dsl dsr > 0:
    // its derivative is w times the (cm, c0, cp) of the selected difference
    double cm = 0.0, c0 = 0.0, cp = 0.0, w = 0.0;
    if (dsl * dsr > 0.0) {
        const double adsl = dabs(dsl), adsr = dabs(dsr), adsc = dabs(dsc);
        const bool inner_r = adsr < adsl;
        const double m = inner_r ? adsr : adsl;
        const double g = dsign1(dsc) * rb;
        if (m < adsc) {
            w = g * dabs_d(inner_r ? dsr : dsl);
            cm = inner_r ? 0.0 : -2.0;
            c0 = inner_r ? -2.0 : 2.0;
            cp = inner_r ? 2.0 : 0.0;
        } else {
            w = g * dabs_d(dsc);
            cm = -0.5;
            cp = 0.5;
        }
    }
    qm = cm != 0.0 ? cm * w : 0.0;
    q0 = c0 != 0.0 ? c0 * w : 0.0;
    qp = cp != 0.0 ? cp * w : 0.0;
}

// B7: face-value adjoints of face p (per direction)
template <bool P2>
__global__ void __launch_bounds__(NTB) k_adj_face(AB A) {
    int i, j, k;
    long long p;
    if (!point(A, i, j, k, p)) return;
    const PerT<P2> P = make_per<P2>(A.n);
    const long long co = (long long)blockIdx.y * A.cs;
    const double* S = A.S + co;
    double bsa[3], bsb[3], bla[3], blb[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) face_bar(A, P, co, S, i, j, k, d, bsa[d], bsb[d], bla[d], blb[d]);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        A.DSA[d][co + p] = bsa[d];
        A.DSB[d][co + p] = bsb[d];
        A.DLA[d][co + p] = bla[d];
        A.DLB[d][co + p] = blb[d];
    }
}

// B8: van Leer slope adjoints of cell p; its slope feeds faces p and p + d
template <bool P2>
__global__ void __launch_bounds__(NTB) k_adj_slope(AB A) {
    int i, j, k;
    long long p;
    if (!point(A, i, j, k, p)) return;
    const PerT<P2> P = make_per<P2>(A.n);
    const long long co = (long long)blockIdx.y * A.cs;
    const double* S = A.S + co;
    double qm[3], q0[3], qp[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double rb = A.DLB[d][co + p] + A.DLA[d][co + P.nb(i, j, k, d, 1)];
        slope_bar(P, S, i, j, k, d, rb, qm[d], q0[d], qp[d]);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        A.QM[d][co + p] = qm[d];
        A.Q0[d][co + p] = q0[d];
        A.QP[d][co + p] = qp[d];
    }
}

// B7 + B8 + B9 in one gather: the state adjoint of cell p with the face
// adjoints of faces p - d .. p + 2d and the slope adjoints of cells
// p - d .. p + d recomputed per direction
template <bool P2>
__global__ void __launch_bounds__(NTB) k_adj_state_fused(AB A) {
    int i, j, k;
    long long p;
    if (!point(A, i, j, k, p)) return;
    const PerT<P2> P = make_per<P2>(A.n);
    const long long co = (long long)blockIdx.y * A.cs;
    const double* S = A.S + co;
    double s = A.G[co + p] + A.S0S[co + p];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int di = d == 0, dj = d == 1, dk = d == 2;
        double fsa[4], fsb[4], fla[4], flb[4];  // faces p + (m - 1) d
#pragma unroll
        for (int m = 0; m < 4; ++m)
            face_bar(A, P, co, S, i + (m - 1) * di, j + (m - 1) * dj, k + (m - 1) * dk, d, fsa[m],
                     fsb[m], fla[m], flb[m]);
        double qm[3], q0[3], qp[3];  // cells p + (m - 1) d
#pragma unroll
        for (int m = 0; m < 3; ++m)
            slope_bar(P, S, i + (m - 1) * di, j + (m - 1) * dj, k + (m - 1) * dk, d,
                      flb[m] + fla[m + 1], qm[m], q0[m], qp[m]);
        s += fsb[1] + fsa[2] + qp[0] + q0[1] + qm[2];
    }
    A.SBF(i, j, k, (int)blockIdx.y) = s;
}

// B9: state adjoint
template <bool P2>
__global__ void __launch_bounds__(NTB) k_adj_state(AB A) {
    int i, j, k;
    long long p;
    if (!point(A, i, j, k, p)) return;
    const PerT<P2> P = make_per<P2>(A.n);
    const long long co = (long long)blockIdx.y * A.cs;
    double s = A.G[co + p] + A.S0S[co + p];
    for (int d = 0; d < 3; ++d) {
        const long long cm = P.nb(i, j, k, d, -1), cp = P.nb(i, j, k, d, 1);
        s += A.DSB[d][co + p] + A.DSA[d][co + cp] + A.QP[d][co + cm] + A.Q0[d][co + p] +
             A.QM[d][co + cp];
    }
    A.SBF(i, j, k, (int)blockIdx.y) = s;
}

// B10: velocity adjoint per face, summed over components in order
// into the caller's face FABs; the duplicate face n of a periodic face FAB
// carries no separate value and is set to zero
template <bool P2, bool STORED>
__global__ void __launch_bounds__(NTB) k_adj_vel(AB A, DA ub0, DA ub1, DA ub2, int has_n) {
    int i, j, k;
    long long p;
    if (!point(A, i, j, k, p)) return;
    const PerT<P2> P = make_per<P2>(A.n);
    double r[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const long long cm = P.nb(i, j, k, d, -1);
        double s = 0.0;
        for (int c = 0; c < A.ncomp; ++c) {
            const long long co = (long long)c * A.cs;
            // (face totals + lower-face trace term) + upper-face trace term of
            // the cell below, in every variant
            double ub;
            if (STORED) {
                ub = A.UBC[d][co + p];
            } else {
                double slb, srb, ubt;
                face_totals(A, P, co, i, j, k, d, true, slb, srb, ubt);
                ub = ubt + A.ULO[d][co + p];
            }
            s += ub + A.UHI[d][co + cm];
        }
        r[d] = s;
    }
    ub0(i, j, k) = r[0];
    ub1(i, j, k) = r[1];
    ub2(i, j, k) = r[2];
    const int n = A.n;
    if (i == 0 && (has_n & 1)) ub0(n, j, k) = 0.0;
    if (j == 0 && (has_n & 2)) ub1(i, n, k) = 0.0;
    if (k == 0 && (has_n & 4)) ub2(i, j, n) = 0.0;
}

// ---------------------------------------------------------------- glue
// periodic ghost fill of a cell FAB (comps) or of a face FAB (nodal in nd,
// face n := face 0), over the FAB's whole extent
struct FillTask {
    DA a;
    int lo[3], hi[3];  // whole FAB
    int nodal;         // -1 cell, else the nodal direction
    int ncomp;
};

__global__ void __launch_bounds__(NTB) k_periodic_fill(const FillTask* __restrict__ tasks, int n) {
    const FillTask& t = tasks[blockIdx.y];
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!region_point(idx, t.lo, t.hi, i, j, k)) return;
    const int q[3] = {i, j, k};
    int s[3];
    bool inside = true;
    for (int d = 0; d < 3; ++d) {
        s[d] = ((q[d] % n) + n) % n;
        if (q[d] != s[d]) inside = false;
    }
    if (inside) return;
    for (int c = 0; c < t.ncomp; ++c) t.a(i, j, k, c) = t.a(s[0], s[1], s[2], c);
}

// dense copy of a FAB's valid cells (or faces 0..N-1) into an N^3 array
__global__ void __launch_bounds__(NTB) k_to_dense(DA a, int comp0, int ncomp, int n, double* out) {
    const long long n3 = (long long)n * n * n;
    const unsigned q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (unsigned)n3) return;
    const long long p = q;
    const unsigned r = q / (unsigned)n;
    const int i = (int)(q - r * (unsigned)n), k = (int)(r / (unsigned)n);
    const int j = (int)(r - (unsigned)k * (unsigned)n);
    for (int c = 0; c < ncomp; ++c) out[(long long)c * n3 + p] = a(i, j, k, comp0 + c);
}

// the update snew = s + dt (-div(u sedge) + 0)
__global__ void __launch_bounds__(NTB) k_step_update(DA s, DA snew, DA u0, DA u1, DA u2, DA e0,
                                                     DA e1, DA e2, int n, int ncomp, double dt,
                                                     double dx0, double dx1, double dx2) {
    const long long n3 = (long long)n * n * n;
    const unsigned q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (unsigned)n3) return;
    const long long p = q;
    const unsigned r = q / (unsigned)n;
    const int i = (int)(q - r * (unsigned)n), k = (int)(r / (unsigned)n);
    const int j = (int)(r - (unsigned)k * (unsigned)n);
    for (int c = 0; c < ncomp; ++c) {
        const double fx1 = u0(i + 1, j, k) * e0(i + 1, j, k, c);
        const double fx0 = u0(i, j, k) * e0(i, j, k, c);
        const double fy1 = u1(i, j + 1, k) * e1(i, j + 1, k, c);
        const double fy0 = u1(i, j, k) * e1(i, j, k, c);
        const double fz1 = u2(i, j, k + 1) * e2(i, j, k + 1, c);
        const double fz0 = u2(i, j, k) * e2(i, j, k, c);
        double divterm = (fx1 - fx0) / dx0;
        divterm += (fy1 - fy0) / dx1;
        divterm += (fz1 - fz0) / dx2;
        snew(i, j, k, c) = s(i, j, k, c) + dt * (-divterm + 0.0);
    }
}

__global__ void __launch_bounds__(NTB) k_fill_zero(double* a, long long n) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) a[p] = 0.0;
}

// the step's preconditions: one periodic box [0, n)^3 at level 0
// ghosts: the step reads s and umac ghosts (>= 3 cells for ppm_type 1, >= 4
// for ppm_type 2, >= 1 face); the
// adjoint reads only valid cells and faces 0..n-1
int step_check(const mx_adv_ctx* c, const mx_fab* s, const mx_fab* u, int ncomp, int& n,
               bool ghosts = true) {
    const mx_adv_params& p = c->p;
    if (p.dim != 3 || (p.ppm_type != 1 && p.ppm_type != 2) || p.ppm_trace_forces != 0)
        return MX_ERR_UNSUPPORTED;
    if (!p.is_periodic[0] || !p.is_periodic[1] || !p.is_periodic[2]) return MX_ERR_UNSUPPORTED;
    if (c->grids[0].size() != 1) return MX_ERR_UNSUPPORTED;
    const mx_box& b = c->grids[0][0];
    n = b.hi[0] - b.lo[0] + 1;
    for (int d = 0; d < 3; ++d) {
        if (b.lo[d] != 0 || b.hi[d] != n - 1) return MX_ERR_UNSUPPORTED;
        if (p.domain_lo[0][d] != 0 || p.domain_hi[0][d] != n - 1) return MX_ERR_UNSUPPORTED;
    }
    if (n < 4 || ncomp < 1 || ncomp > MX_MAX_SCAL) return MX_ERR_ARG;
    if ((long long)n * n * n * ncomp > 0x7fffffffLL) return MX_ERR_UNSUPPORTED;
    if (!s || !u) return MX_ERR_ARG;
    HBox v = hbox(b);
    // the edge-state stencils of the step reach 3 (ppm_type 1) or 4
    // (ppm_type 2) cells past the box
    const int g = p.ppm_type == 2 ? 4 : 3;
    if (!covers(*s, ghosts ? grow(v, 3, g) : v, ncomp)) return MX_ERR_BOX;
    for (int d = 0; d < 3; ++d)
        if (!covers(u[d], ghosts ? grow(nodal(v, d), 3, 1) : v, 1)) return MX_ERR_BOX;
    return MX_OK;
}

// is the FAB exactly a dense (ncomp, n, n, n) array starting at cell 0?
bool dense_fab(const mx_fab& f, int n, int ncomp) {
    const long long n2 = (long long)n * n;
    return f.begin[0] == 0 && f.begin[1] == 0 && f.begin[2] == 0 && f.jstride == n &&
           f.kstride == n2 && (ncomp == 1 || f.nstride == n2 * n);
}

dim3 grid1(long long npts) { return dim3((unsigned)((npts + NTB - 1) / NTB)); }

int periodic_fill(mx_adv_ctx* c, const mx_fab* s, const mx_fab* u, int ncomp, int n) {
    std::vector<FillTask> tasks;
    long long maxpts = 0;
    auto add = [&](const mx_fab& f, int nodal_dir, int nc) {
        FillTask t;
        t.a = to_da(f);
        for (int d = 0; d < 3; ++d) {
            t.lo[d] = f.begin[d];
            t.hi[d] = f.end[d] - 1;
        }
        t.nodal = nodal_dir;
        t.ncomp = nc;
        long long np = 1;
        for (int d = 0; d < 3; ++d) np *= (long long)(t.hi[d] - t.lo[d] + 1);
        maxpts = std::max(maxpts, np);
        tasks.push_back(t);
    };
    add(*s, -1, ncomp);
    for (int d = 0; d < 3; ++d) add(u[d], d, 1);
    Table tab(c);
    size_t off = tab.add(tasks);
    int rc = tab.upload();
    if (rc) return rc;
    dim3 gr = grid_for(maxpts, (int)tasks.size(), 1, NTB);
    MX_LAUNCH(c, "k_periodic_fill", k_periodic_fill, gr, dim3(NTB), tab.dev<FillTask>(off), n);
    return MX_OK;
}

// the tape and reverse kernels of the step adjoint
template <bool P2>
int vjp_kernels(mx_adv_ctx* c, const AB& A, dim3 g1, dim3 gc, mx_fab* ubar) {
    const bool ppm2 = c->p.ppm_type == 2;
    if (ppm2)
        MX_LAUNCH(c, "k_tape_ppm", (k_tape_ppm<2, P2>), gc, dim3(NTB), A);
    else
        MX_LAUNCH(c, "k_tape_ppm", (k_tape_ppm<1, P2>), gc, dim3(NTB), A);
    MX_LAUNCH(c, "k_tape_corner", k_tape_corner<P2>, gc, dim3(NTB), A);
    MX_LAUNCH(c, "k_tape_edge", k_tape_edge<P2>, gc, dim3(NTB), A);
    // the transverse-term and trace gathers run faster with compare
    // indexing on MI355X (measured), whatever n is
    MX_LAUNCH(c, "k_adj_tterm", k_adj_tterm<false>, gc, dim3(NTB), A);
    if (ppm2) {
        // ppm_type 2: the Colella-Sekora trace and state sweeps (the kernel
        // arrangement options apply to ppm_type 1)
        MX_LAUNCH(c, "k_adj_trace", k_adj_trace2<false>, gc, dim3(NTB), A);
        MX_LAUNCH(c, "k_adj_state", k_adj_state2<P2>, gc, dim3(NTB), A);
    } else if (c->adj_face_sums) {
        MX_LAUNCH(c, "k_adj_fsum", k_adj_fsum<P2>, gc, dim3(NTB), A);
        MX_LAUNCH(c, "k_adj_trace", (k_adj_trace<false, true, false>), gc, dim3(NTB), A);
    } else if (c->adj_trace_rolled) {
        MX_LAUNCH(c, "k_adj_trace", (k_adj_trace<false, false, true>), gc, dim3(NTB), A);
    } else {
        MX_LAUNCH(c, "k_adj_trace", (k_adj_trace<false, false, false>), gc, dim3(NTB), A);
    }
    if (ppm2) {
        // the state sweep ran above
    } else if (c->adj_state_fused) {
        MX_LAUNCH(c, "k_adj_state_fused", k_adj_state_fused<P2>, gc, dim3(NTB), A);
    } else {
        MX_LAUNCH(c, "k_adj_face", k_adj_face<P2>, gc, dim3(NTB), A);
        MX_LAUNCH(c, "k_adj_slope", k_adj_slope<P2>, gc, dim3(NTB), A);
        MX_LAUNCH(c, "k_adj_state", k_adj_state<P2>, gc, dim3(NTB), A);
    }
    // velocity adjoint, into the caller's face FABs
    int has_n = 0;
    for (int d = 0; d < 3; ++d)
        if (ubar[d].end[d] > A.n) has_n |= 1 << d;
    if (c->adj_face_sums && !ppm2)
        MX_LAUNCH(c, "k_adj_vel", (k_adj_vel<P2, false>), g1, dim3(NTB), A, to_da(ubar[0]),
                  to_da(ubar[1]), to_da(ubar[2]), has_n);
    else
        MX_LAUNCH(c, "k_adj_vel", (k_adj_vel<P2, true>), g1, dim3(NTB), A, to_da(ubar[0]),
                  to_da(ubar[1]), to_da(ubar[2]), has_n);
    return MX_OK;
}

}  // namespace

extern "C" {

int mx_adv_scal_step(mx_adv_ctx* c, mx_fab* s, mx_fab* umac, mx_fab* snew, int ncomp) {
    if (!c || !s || !umac || !snew) return MX_ERR_ARG;
    int n = 0;
    int rc = step_check(c, s, umac, ncomp, n);
    if (rc) return rc;
    HBox v = hbox(c->grids[0][0]);
    if (!covers(*snew, v, ncomp)) return MX_ERR_BOX;
    if ((rc = periodic_fill(c, s, umac, ncomp, n))) return rc;
    // edge states into workspace face FABs (no ghosts), zero force with 1 ghost
    const long long fn = (long long)(n + 1) * n * n;
    const long long cn = (long long)(n + 2) * (n + 2) * (n + 2);
    if ((rc = ws_reserve(c, (3 * fn * ncomp + cn * ncomp) * sizeof(double) + 8192))) return rc;
    mx_fab se[3], fo;
    for (int d = 0; d < 3; ++d) {
        se[d].p = ws_take(c, (size_t)(fn * ncomp));
        const int ex[3] = {n + (d == 0), n + (d == 1), n + (d == 2)};
        se[d].jstride = ex[0];
        se[d].kstride = (long long)ex[0] * ex[1];
        se[d].nstride = fn;
        for (int q = 0; q < 3; ++q) {
            se[d].begin[q] = 0;
            se[d].end[q] = ex[q];
        }
        se[d].ncomp = ncomp;
    }
    fo.p = ws_take(c, (size_t)(cn * ncomp));
    fo.jstride = n + 2;
    fo.kstride = (long long)(n + 2) * (n + 2);
    fo.nstride = cn;
    for (int q = 0; q < 3; ++q) {
        fo.begin[q] = -1;
        fo.end[q] = n + 1;
    }
    fo.ncomp = ncomp;
    MX_LAUNCH(c, "k_fill_zero", k_fill_zero, grid1(cn * ncomp), dim3(NTB), fo.p, cn * ncomp);
    mx_multifab sm{1, s}, fm{1, &fo};
    mx_face_multifab sem, um;
    for (int d = 0; d < 3; ++d) {
        sem.dir[d] = {1, &se[d]};
        um.dir[d] = {1, &umac[d]};
    }
    std::vector<mx_bcrec> bcs(ncomp);
    for (auto& b : bcs)
        for (int d = 0; d < 3; ++d) b.lo[d] = b.hi[d] = MX_BC_INT_DIR;
    const int saved = c->fused_edge;
    c->fused_edge = 1;
    rc = make_edge_scal(c, 0, &sm, &sem, &um, &fm, 0, bcs.data(), ncomp, 0, 0, ncomp, 0);
    c->fused_edge = saved;
    if (rc) return rc;
    MX_LAUNCH(c, "k_step_update", k_step_update, grid1((long long)n * n * n), dim3(NTB), to_da(*s),
              to_da(*snew), to_da(umac[0]), to_da(umac[1]), to_da(umac[2]), to_da(se[0]),
              to_da(se[1]), to_da(se[2]), n, ncomp, c->p.dt, c->p.dx[0][0], c->p.dx[0][1],
              c->p.dx[0][2]);
    return MX_OK;
}

int mx_adv_scal_step_vjp(mx_adv_ctx* c, mx_fab* s, mx_fab* umac, const mx_fab* gbar,
                         mx_fab* sbar, mx_fab* ubar, int ncomp) {
    if (!c || !s || !umac || !gbar || !sbar || !ubar) return MX_ERR_ARG;
    int n = 0;
    int rc = step_check(c, s, umac, ncomp, n, false);
    if (rc) return rc;
    HBox v = hbox(c->grids[0][0]);
    if (!covers(*gbar, v, ncomp) || !covers(*sbar, v, ncomp)) return MX_ERR_BOX;
    for (int d = 0; d < 3; ++d)
        if (!covers(ubar[d], v, 1)) return MX_ERR_BOX;
    const long long n3 = (long long)n * n * n;
    // dense arrays: S and G (ncomp each), 63 per-component intermediates and
    // adjoints, U (3)
    const size_t per = (size_t)n3 * sizeof(double) + 256;
    const bool fs = c->adj_face_sums && c->p.ppm_type == 1;
    const size_t count = (size_t)ncomp * (63 + (fs ? 9 : 0) + (c->p.ppm_type == 2 ? 12 : 0)) + 3;
    if ((rc = ws_reserve(c, per * count + 4096))) return rc;
    AB A;
    std::memset(&A, 0, sizeof(A));
    auto take = [&](int m) { return ws_take(c, (size_t)n3 * m); };
    // inputs that are already dense (ncomp, n, n, n) arrays are read in place
    const bool sd = dense_fab(*s, n, ncomp), gd = dense_fab(*gbar, n, ncomp);
    bool ud[3];
    for (int d = 0; d < 3; ++d) ud[d] = dense_fab(umac[d], n, 1);
    double* S = sd ? s->p : take(ncomp);
    double* G = gd ? gbar->p : take(ncomp);
    for (int d = 0; d < 3; ++d) A.U[d] = ud[d] ? umac[d].p : take(1);
    for (int d = 0; d < 3; ++d) {
        A.IP[d] = take(ncomp);
        A.IM[d] = take(ncomp);
        A.SIMH[d] = take(ncomp);
        A.SE[d] = take(ncomp);
        A.TB[d] = take(ncomp);
        A.UBC[d] = take(ncomp);
        if (c->p.ppm_type == 2)
            for (int o = 0; o < 4; ++o) A.DV[d][o] = take(ncomp);
        if (fs) {
            A.SLB[d] = take(ncomp);
            A.SRB[d] = take(ncomp);
        }
        A.FLB[d] = take(ncomp);
        A.FRB[d] = take(ncomp);
        A.UHI[d] = take(ncomp);
        if (fs) A.ULO[d] = take(ncomp);
        A.DSA[d] = take(ncomp);
        A.DSB[d] = take(ncomp);
        A.DLA[d] = take(ncomp);
        A.DLB[d] = take(ncomp);
        A.QM[d] = take(ncomp);
        A.Q0[d] = take(ncomp);
        A.QP[d] = take(ncomp);
    }
    for (int q = 0; q < 6; ++q) {
        A.COR[q] = take(ncomp);
        A.EB[q] = take(ncomp);
    }
    A.S = S;
    A.G = G;
    A.S0S = take(ncomp);
    A.SBF = to_da(*sbar);
    A.cs = n3;
    A.n = n;
    A.ncomp = ncomp;
    A.dt = c->p.dt;
    A.rel_eps = c->p.rel_eps;
    for (int d = 0; d < 3; ++d) {
        A.dx[d] = c->p.dx[0][d];
        A.fdt[d] = A.dt / A.dx[d];
        const double dt4 = 0.25 * A.dt, dt6 = A.dt / 6.0;
        A.f4[d] = dt4 / A.dx[d];
        A.f6[d] = dt6 / A.dx[d];
    }

    const dim3 g1 = grid1(n3);
    const dim3 gc(g1.x, (unsigned)ncomp);
    if (!sd) MX_LAUNCH(c, "k_to_dense", k_to_dense, g1, dim3(NTB), to_da(*s), 0, ncomp, n, S);
    if (!gd)
        MX_LAUNCH(c, "k_to_dense", k_to_dense, g1, dim3(NTB), to_da(*gbar), 0, ncomp, n, G);
    for (int d = 0; d < 3; ++d)
        if (!ud[d])
            MX_LAUNCH(c, "k_to_dense", k_to_dense, g1, dim3(NTB), to_da(umac[d]), 0, 1, n,
                      const_cast<double*>(A.U[d]));
    const bool p2 = c->adj_pow2 && (n & (n - 1)) == 0;
    return p2 ? vjp_kernels<true>(c, A, g1, gc, ubar) : vjp_kernels<false>(c, A, g1, gc, ubar);
}

}  // extern "C"
