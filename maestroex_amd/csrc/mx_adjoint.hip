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
template <int D, class PT>
__device__ __forceinline__ void ppm_cell(const AB& A, const PT& P, const double* S, int i, int j,
                                         int k, int pos, double& Ip, double& Im) {
    double v[9];
#pragma unroll
    for (int o = -4; o <= 4; ++o) v[4 + o] = (o >= -2 && o <= 2) ? S[P.nb(i, j, k, D, o)] : 0.0;
    double sm, sp;
    ppm_edges<1, D>(v, pos, 0, A.n - 1, MX_BC_INT_DIR, MX_BC_INT_DIR, sm, sp);
    ppm_trace(v[4], sm, sp, A.U[D][P.nb(i, j, k, D, 1)], A.U[D][P.at(i, j, k)], A.dt, A.dx[D],
              A.rel_eps, Ip, Im);
}

template <int D, class PT>
__device__ __forceinline__ void tape_dir(const AB& A, const PT& P, const double* S, int i, int j,
                                         int k, long long p, double& ip, double& im, double& h) {
    const int pos = D == 0 ? i : (D == 1 ? j : k);
    ppm_cell<D>(A, P, S, i, j, k, pos, ip, im);
    double ipm, imm;
    ppm_cell<D>(A, P, S, i - (D == 0), j - (D == 1), k - (D == 2), pos - 1, ipm, imm);
    h = upwind(ipm, im, A.U[D][p], 0.0);
}

template <bool P2>
__global__ void __launch_bounds__(NTB) k_tape_ppm(AB A) {
    int i, j, k;
    long long p;
    if (!point(A, i, j, k, p)) return;
    const PerT<P2> P = make_per<P2>(A.n);
    const long long co = (long long)blockIdx.y * A.cs;
    const double* S = A.S + co;
    double ip[3], imv[3], h[3];
    tape_dir<0>(A, P, S, i, j, k, p, ip[0], imv[0], h[0]);
    tape_dir<1>(A, P, S, i, j, k, p, ip[1], imv[1], h[1]);
    tape_dir<2>(A, P, S, i, j, k, p, ip[2], imv[2], h[2]);
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
        double smb = 0.0, spb = 0.0, s0b = 0.0, uhib = 0.0, ulob = 0.0;
        if (uhi > A.rel_eps) {
            const double sg = dabs(uhi) * A.dt / h;
            const double f = 1.0 - C23 * sg;
            const double Aq = sp - sm - f * s6;
            spb += Ipb * (1.0 - 0.5 * sg * (1.0 + 3.0 * f));
            smb += Ipb * (-0.5 * sg * (-1.0 + 3.0 * f));
            s0b += Ipb * (3.0 * sg * f);
            const double sgb = Ipb * (-0.5 * Aq - 0.5 * sg * (C23 * s6));
            uhib += sgb * dabs_d(uhi) * A.dt / h;
        } else {
            s0b += Ipb;
        }
        if (ulo < -A.rel_eps) {
            const double sg = dabs(ulo) * A.dt / h;
            const double f = 1.0 - C23 * sg;
            const double Bq = sp - sm + f * s6;
            smb += Imb * (1.0 + 0.5 * sg * (-1.0 - 3.0 * f));
            spb += Imb * (0.5 * sg * (1.0 - 3.0 * f));
            s0b += Imb * (3.0 * sg * f);
            const double sgb = Imb * (0.5 * Bq - 0.5 * sg * (C23 * s6));
            ulob += sgb * dabs_d(ulo) * A.dt / h;
        } else {
            s0b += Imb;
        }
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
// ghosts: the step reads s and umac ghosts (>= 3 cells, >= 1 face); the
// adjoint reads only valid cells and faces 0..n-1
int step_check(const mx_adv_ctx* c, const mx_fab* s, const mx_fab* u, int ncomp, int& n,
               bool ghosts = true) {
    const mx_adv_params& p = c->p;
    if (p.dim != 3 || p.ppm_type != 1 || p.ppm_trace_forces != 0) return MX_ERR_UNSUPPORTED;
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
    if (!covers(*s, ghosts ? grow(v, 3, 3) : v, ncomp)) return MX_ERR_BOX;
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
    MX_LAUNCH(c, "k_tape_ppm", k_tape_ppm<P2>, gc, dim3(NTB), A);
    MX_LAUNCH(c, "k_tape_corner", k_tape_corner<P2>, gc, dim3(NTB), A);
    MX_LAUNCH(c, "k_tape_edge", k_tape_edge<P2>, gc, dim3(NTB), A);
    // the transverse-term and trace gathers run faster with compare
    // indexing on MI355X (measured), whatever n is
    MX_LAUNCH(c, "k_adj_tterm", k_adj_tterm<false>, gc, dim3(NTB), A);
    if (c->adj_face_sums) {
        MX_LAUNCH(c, "k_adj_fsum", k_adj_fsum<P2>, gc, dim3(NTB), A);
        MX_LAUNCH(c, "k_adj_trace", (k_adj_trace<false, true, false>), gc, dim3(NTB), A);
    } else if (c->adj_trace_rolled) {
        MX_LAUNCH(c, "k_adj_trace", (k_adj_trace<false, false, true>), gc, dim3(NTB), A);
    } else {
        MX_LAUNCH(c, "k_adj_trace", (k_adj_trace<false, false, false>), gc, dim3(NTB), A);
    }
    if (c->adj_state_fused) {
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
    if (c->adj_face_sums)
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
    const size_t count = (size_t)ncomp * (63 + (c->adj_face_sums ? 9 : 0)) + 3;
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
        if (c->adj_face_sums) {
            A.SLB[d] = take(ncomp);
            A.SRB[d] = take(ncomp);
        }
        A.FLB[d] = take(ncomp);
        A.FRB[d] = take(ncomp);
        A.UHI[d] = take(ncomp);
        if (c->adj_face_sums) A.ULO[d] = take(ncomp);
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
