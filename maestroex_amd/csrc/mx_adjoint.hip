// mx_adjoint.hip -- a differentiable scalar advection step and its
// hand-written adjoint (vector-Jacobian product).
//
// The step, on one fully periodic box [0, N)^3 or over a box list with
// periodic or outflow sides, with a fixed MAC velocity, for every scalar
// component (the enthalpy step of EnthalpyAdvance, predict_rhohprime, for
// (rho h)' with a zero base state; with is_conservative the conservative
// prediction, as the reference's predict_rhoX species and predict_rhoh
// enthalpy steps use it):
//   sedge = MakeEdgeScal(s, umac, force = 0, advective form)  (ppm_type 1, 2;
//                                      MaestroEnthalpyAdvance.cpp:143-147)
//   F     = umac * (0 + sedge)        (MakeRhoHFlux, MaestroMakeFlux.cpp:699-708)
//   snew  = s + dt * (-div F + 0)     (UpdateScal, MaestroAdvection.cpp:120-135)
// The forward runs the production kernels (ghost fill, the fused edge-state
// kernel) plus an update kernel.
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

struct AB;

// periodic indexing of a dense N^3 array (N^3 fits 31 bits; checked on the
// host, so all index arithmetic is 32-bit).  P2: N is a power of two and a
// coordinate within one period of [0, N) wraps with a mask.
template <bool P2>
struct PerT {
    static constexpr bool kGen = false;
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
    // no domain sides: every face is interior, every point a cell
    __device__ __forceinline__ int bnd(int, int) const { return 0; }
    __device__ __forceinline__ bool cell(int, int, int) const { return true; }
    __device__ static PerT make(const AB& A);
};

template <bool P2>
__device__ __forceinline__ PerT<P2> make_per(int n) {
    PerT<P2> P;
    P.n = n;
    P.m = n - 1;
    P.lg = 31 - __clz(n);
    return P;
}

// The general layout: nn[d] cells per direction, each direction periodic
// (positions wrap) or outflow.  Along an outflow direction the kernels visit
// positions 0 .. nn[d] (cells 0 .. nn[d]-1 and face nn[d]; the values at the
// position-nn[d] cell are unused), stored after pad[d] points of padding on
// each side that hold the state's foextrap copies and zero in every other
// array, so that reads up to pad[d] points outside need no test.
struct GenT {
    static constexpr bool kGen = true;
    // one scalar per direction (no arrays: an array indexed at run time goes
    // to scratch memory)
    int nn0, nn1, nn2, per0, per1, per2, pad0, pad1, pad2;
    int sj, sk;
    // the value of direction d by masks (a select chain is turned into a
    // load from a scratch copy of the struct at a selected offset)
    __device__ __forceinline__ static int pick(int d, int a, int b, int c) {
        return (a & -(int)(d == 0)) | (b & -(int)(d == 1)) | (c & -(int)(d == 2));
    }
    __device__ __forceinline__ int nn_of(int d) const { return pick(d, nn0, nn1, nn2); }
    __device__ __forceinline__ int per_of(int d) const { return pick(d, per0, per1, per2); }
    __device__ __forceinline__ static int wrap(int x, int per, int n, int pad) {
        if (per) return x < 0 ? x + n : (x >= n ? x - n : x);
        return x + pad;
    }
    __device__ __forceinline__ int at(int i, int j, int k) const {
        return wrap(i, per0, nn0, pad0) + sj * wrap(j, per1, nn1, pad1) +
               sk * wrap(k, per2, nn2, pad2);
    }
    __device__ __forceinline__ int nb(int i, int j, int k, int d, int o) const {
        return at(i + (d == 0 ? o : 0), j + (d == 1 ? o : 0), k + (d == 2 ? o : 0));
    }
    // 1: a d-face at the low outflow side, 2: at the high one, else 0
    __device__ __forceinline__ int bnd(int d, int pos) const {
        if (per_of(d)) return 0;
        return pos == 0 ? 1 : (pos == nn_of(d) ? 2 : 0);
    }
    __device__ __forceinline__ bool cell(int i, int j, int k) const {
        return i < nn0 && j < nn1 && k < nn2;
    }
    __device__ static GenT make(const AB& A);
};

__host__ __device__ __forceinline__ int posd(int d, int i, int j, int k) {
    return d == 0 ? i : (d == 1 ? j : k);
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
    long long cs;      // stored points per component (N^3 when periodic)
    long long npts;    // visited points: ext[0] ext[1] ext[2]
    int ext[3];
    int n, ncomp;
    GenT gen;          // the general layout (outflow sides)
    int xcd;           // the grid's x is a multiple of 8 and each XCD walks one
                       // contiguous eighth of it
    int fold;          // ghost reach of the state adjoint (general layout)
    double dt, rel_eps, dx[3];
    // per-direction factors, evaluated once on the host by the same
    // expressions: dt/h, (dt/4)/h (edge terms), (dt/6)/h (transverse terms);
    // the conservative form's (dt/2)/h (edge terms) and (dt/3)/h (transverse
    // terms), dt/3 and dx[0] (MakeDivU's divisor)
    double fdt[3], f4[3], f6[3];
    double f2[3], f3[3], dt3, dx0;
};

template <bool P2>
__device__ __forceinline__ PerT<P2> PerT<P2>::make(const AB& A) {
    return make_per<P2>(A.n);
}
__device__ __forceinline__ GenT GenT::make(const AB& A) { return A.gen; }

// the visited point of this thread: its position (i,j,k) and stored index p
template <class PT>
__device__ __forceinline__ bool point(const AB& A, const PT& P, int& i, int& j, int& k,
                                      long long& p) {
    // workgroups go to the 8 XCDs round robin by linear id: with A.xcd the
    // block of XCD x takes the x-th eighth of the points, in order, so that
    // stencil neighbours stay in that XCD's L2
    unsigned b = blockIdx.x;
    if (A.xcd) b = (b & 7u) * (gridDim.x >> 3) + (b >> 3);
    const unsigned q = b * blockDim.x + threadIdx.x;
    if (q >= (unsigned)A.npts) return false;
    const unsigned e0 = (unsigned)A.ext[0], e1 = (unsigned)A.ext[1];
    const unsigned r = q / e0;
    i = (int)(q - r * e0);
    k = (int)(r / e1);
    j = (int)(r - (unsigned)k * e1);
    p = PT::kGen ? (long long)P.at(i, j, k) : (long long)q;
    return true;
}

// the foextrap rule at an outflow side (1: low, 2: high) for the predictor,
// corner-coupled and edge states (MaestroMakeEdgeScalars.cpp:709-738,
// :931-961, :1446-1480): the outside state takes the inside one
__device__ __forceinline__ double upwind_b(double l, double r, double u, double thr, int b);
// (selects, not branches: assignments under branches to the two
// references put them in scratch memory)
__device__ __forceinline__ void bfix(int b, double& wl, double& wr) {
    const double sum = wl + wr;
    const double l = b == 1 ? 0.0 : (b == 2 ? sum : wl);
    const double r = b == 1 ? sum : (b == 2 ? 0.0 : wr);
    wl = l;
    wr = r;
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

__device__ __forceinline__ double upwind_b(double l, double r, double u, double thr, int b) {
    const double L = b == 1 ? r : l;
    const double R = b == 2 ? l : r;
    return upwind(L, R, u, thr);
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
    h = upwind_b(ipm, im, A.U[D][p], 0.0, P.bnd(D, pos));
}

template <int PPM, class PT>
__global__ void __launch_bounds__(NTB) k_tape_ppm(AB A) {
    const PT P = PT::make(A);
    int i, j, k;
    long long p;
    if (!point(A, P, i, j, k, p)) return;
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

// MakeDivU of cell (i,j,k): the six face differences over dx[0]
// (MaestroMakeEdgeScalars.cpp:655-658)
template <class PT>
__device__ __forceinline__ double divu(const AB& A, const PT& P, int i, int j, int k) {
    const long long c = P.at(i, j, k);
    double dv = A.U[0][P.nb(i, j, k, 0, 1)] - A.U[0][c] + A.U[1][P.nb(i, j, k, 1, 1)] - A.U[1][c] +
                A.U[2][P.nb(i, j, k, 2, 1)] - A.U[2][c];
    return dv / A.dx0;
}

// the state `base` of cell c corrected by its transverse term in direction
// t: advective base - dt/6h_t (u+ + u-) (simh+ - simh-); conservative
// base - dt/3h_t (simh+ u+ - simh- u-) - dt/3 s divu + dt/3h_t s (u+ - u-)
// (MaestroMakeEdgeScalars.cpp:906-929)
template <bool CONS, class PT>
__device__ __forceinline__ double tcorr(const AB& A, const PT& P, long long co, double base,
                                        int i, int j, int k, int t) {
    const long long c = P.at(i, j, k);
    const long long c1 = P.nb(i, j, k, t, 1);
    const double* H = A.SIMH[t] + co;
    if (!CONS) return base - A.f6[t] * (A.U[t][c1] + A.U[t][c]) * (H[c1] - H[c]);
    const double sc = A.S[co + c];
    return base - A.f3[t] * (H[c1] * A.U[t][c1] - H[c] * A.U[t][c]) -
           A.dt3 * sc * divu(A, P, i, j, k) + A.f3[t] * sc * (A.U[t][c1] - A.U[t][c]);
}

// A3: corner-coupled states on d-faces, corrected by t
template <class PT, bool CONS>
__global__ void __launch_bounds__(NTB) k_tape_corner(AB A) {
    const PT P = PT::make(A);
    int i, j, k;
    long long p;
    if (!point(A, P, i, j, k, p)) return;
    const long long co = (long long)blockIdx.y * A.cs;
    double cor[6];
    for (int d = 0; d < 3; ++d) {
        const int im = i - (d == 0), jm = j - (d == 1), km = k - (d == 2);
        for (int t = 0; t < 3; ++t) {
            if (t == d) continue;
            double l = A.IP[d][co + P.at(im, jm, km)];
            double r = A.IM[d][co + p];
            l = tcorr<CONS>(A, P, co, l, im, jm, km, t);
            r = tcorr<CONS>(A, P, co, r, i, j, k, t);
            cor[pidx(d, t)] = upwind_b(l, r, A.U[d][p], A.rel_eps, P.bnd(d, posd(d, i, j, k)));
        }
    }
    for (int q = 0; q < 6; ++q) A.COR[q][co + p] = cor[q];
}

// edge term of cell c for the corner state on t-faces corrected by o:
// advective dt/4h_t (u+ + u-) (K+ - K-), conservative dt/2h_t (K+ u+ - K- u-)
// (MaestroMakeEdgeScalars.cpp:1411-1438)
template <bool CONS, class PT>
__device__ __forceinline__ double et(const AB& A, const PT& P, long long co, int i, int j, int k,
                                     int t, int o) {
    const long long c = P.at(i, j, k);
    const long long c1 = P.nb(i, j, k, t, 1);
    const double* K = A.COR[pidx(t, o)] + co;
    if (CONS) return A.f2[t] * (K[c1] * A.U[t][c1] - K[c] * A.U[t][c]);
    return A.f4[t] * (A.U[t][c1] + A.U[t][c]) * (K[c1] - K[c]);
}

// the conservative form's normal term of cell c on its d-faces,
// dt/2h_d s (u_d+ - u_d-)
template <class PT>
__device__ __forceinline__ double nterm(const AB& A, const PT& P, long long co, int i, int j,
                                        int k, int d) {
    const long long c = P.at(i, j, k);
    return A.f2[d] * A.S[co + c] * (A.U[d][P.nb(i, j, k, d, 1)] - A.U[d][c]);
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
    bfix(P.bnd(d, posd(d, i, j, k)), wl, wr);
    return Fb * u;
}

// A4: edge states
template <class PT, bool CONS>
__global__ void __launch_bounds__(NTB) k_tape_edge(AB A) {
    const PT P = PT::make(A);
    int i, j, k;
    long long p;
    if (!point(A, P, i, j, k, p)) return;
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
            l = l - et<CONS>(A, P, co, im, jm, km, t, o);
            r = r - et<CONS>(A, P, co, i, j, k, t, o);
        }
        if (CONS) {
            l = l - nterm(A, P, co, im, jm, km, d);
            r = r - nterm(A, P, co, i, j, k, d);
        }
        const double sedgel = l + dt2 * 0.0;
        const double sedger = r + dt2 * 0.0;
        se[d] = upwind_b(sedgel, sedger, A.U[d][p], A.rel_eps, P.bnd(d, posd(d, i, j, k)));
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
// transverse term (predictor upwinding).  The conservative form's flux
// differences give the corner and upwind states the weights dt/2h u0 and
// dt/3h u0, and its velocity adjoint also collects the state terms of the
// transverse terms (s (u+ - u-) and s divu of both cells) and of the normal
// edge terms (s (u+ - u-))
template <bool CONS, class PT>
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
    const int bt = P.bnd(t, posd(t, i, j, k));
    bfix(bt, wl, wr);
    ub = want_u ? Fb * A.SE[t][co + f] : 0.0;
    slb = wl * seb;
    srb = wr * seb;
    // corner states on t-faces, through the edge terms of the cells below
    // and above the face
    const double b = CONS ? A.f2[t] : A.f4[t];
    double ubc = 0.0, slbc = 0.0, srbc = 0.0;
#pragma unroll
    for (int o = 0; o < 3; ++o) {
        if (o == t) continue;
        const int q = pidx(t, o);
        const double ebm = A.EB[q][co + fm];
        const double eb0 = A.EB[q][co + f];
        const double kb = CONS ? ebm * b * u0 - eb0 * b * u0
                               : ebm * b * (u0 + um) - eb0 * b * (up + u0);
        if (want_u) {
            const double* K = A.COR[q] + co;
            if (CONS)
                ubc += ebm * b * K[f] - eb0 * b * K[f];
            else
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
    const double a = CONS ? A.f3[t] : A.f6[t];
    const double tbm = A.TB[t][co + fm], tb0 = A.TB[t][co + f];
    const double hb = CONS ? tbm * a * u0 - tb0 * a * u0
                           : tbm * a * (u0 + um) - tb0 * a * (up + u0);
    if (want_u) {
        const double* H = A.SIMH[t] + co;
        if (!CONS) {
            ub = ub + (tbm * a * (H[f] - H[fm]) + tb0 * a * (H[fp] - H[f]));
        } else {
            // the cells below (fm) and above (f) the face: their transverse
            // terms in t (simh u and s (u+ - u-)), the divergence in all
            // three of their transverse terms, and their normal edge terms
            // in t
            const double sm = A.S[co + fm], s0 = A.S[co + f];
            const double tsm = A.TB[0][co + fm] + A.TB[1][co + fm] + A.TB[2][co + fm];
            const double ts0 = A.TB[0][co + f] + A.TB[1][co + f] + A.TB[2][co + f];
            const int qn = t == 0 ? pidx(1, 2) : (t == 1 ? pidx(0, 2) : pidx(0, 1));
            const double enm = A.EB[qn][co + fm], en0 = A.EB[qn][co + f];
            ub = ub + (tbm * a * H[f] - tb0 * a * H[f]);
            ub = ub + (tb0 * a * s0 - tbm * a * sm);
            ub = ub + (A.dt3 / A.dx0) * (sm * tsm - s0 * ts0);
            ub = ub + A.f2[t] * (enm * sm - en0 * s0);
        }
    }
    double wl0, wr0;
    weights(u0, 0.0, wl0, wr0);
    bfix(bt, wl0, wr0);
    slb = slb + wl0 * hb;
    srb = srb + wr0 * hb;
}

// adjoint of the corner state q on d-face (i,j,k) (through the edge terms of
// the cells on both sides of it), with its upwinding weights: the left and
// right predictor-state adjoints it passes on
template <bool CONS, class PT>
__device__ __forceinline__ void corner_bar(const AB& A, const PT& P, long long co, int i, int j,
                                           int k, int d, int q, double& lb, double& rb) {
    const double b = CONS ? A.f2[d] : A.f4[d];
    const long long c = P.at(i, j, k);
    const long long fm = P.nb(i, j, k, d, -1), fp = P.nb(i, j, k, d, 1);
    const double u0 = A.U[d][c], um = A.U[d][fm], up = A.U[d][fp];
    const double ebm = A.EB[q][co + fm];
    const double eb0 = A.EB[q][co + c];
    const double kb = CONS ? ebm * b * u0 - eb0 * b * u0
                           : ebm * b * (u0 + um) - eb0 * b * (up + u0);
    double wl, wr;
    weights(u0, A.rel_eps, wl, wr);
    bfix(P.bnd(d, posd(d, i, j, k)), wl, wr);
    lb = wl * kb;
    rb = wr * kb;
}

// the conservative form's direct terms of the state of cell (i,j,k): its
// transverse terms' dt/3 s divu - dt/3h_t s (u+ - u-) and its normal edge
// terms' dt/2h_d s (u+ - u-), with their adjoints TB and EB
template <class PT>
__device__ __forceinline__ double cons_direct(const AB& A, const PT& P, long long co, int i,
                                              int j, int k) {
    const long long c = P.at(i, j, k);
    const double dv = divu(A, P, i, j, k);
    double r = 0.0;
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        const double du = A.U[t][P.nb(i, j, k, t, 1)] - A.U[t][c];
        const int qn = t == 0 ? pidx(1, 2) : (t == 1 ? pidx(0, 2) : pidx(0, 1));
        r += A.TB[t][co + c] * (A.dt3 * dv - A.f3[t] * du) + A.EB[qn][co + c] * (A.f2[t] * du);
    }
    return r;
}

// B4: per-cell transverse-term adjoints: the transverse term of cell p in
// direction o enters the corner states on the d-faces p (right state) and
// p + d (left state), d != o
template <class PT, bool CONS>
__global__ void __launch_bounds__(NTB) k_adj_tterm(AB A) {
    const PT P = PT::make(A);
    int i, j, k;
    long long p;
    if (!point(A, P, i, j, k, p)) return;
    const long long co = (long long)blockIdx.y * A.cs;
    double tb[3];
    for (int o = 0; o < 3; ++o) {
        double s = 0.0;
        for (int d = 0; d < 3; ++d) {
            if (d == o) continue;
            const int q = pidx(d, o);
            double lb1, rb1, lb0, rb0;
            corner_bar<CONS>(A, P, co, i + (d == 0), j + (d == 1), k + (d == 2), d, q, lb1, rb1);
            corner_bar<CONS>(A, P, co, i, j, k, d, q, lb0, rb0);
            s += lb1 + rb0;
        }
        tb[o] = -s;
    }
    for (int o = 0; o < 3; ++o) A.TB[o][co + p] = tb[o];
}

// B3 + B5 stored once per face: the left and right predictor-state adjoints
// of every d-face (for the trace variant that reads them)
template <class PT, bool CONS>
__global__ void __launch_bounds__(NTB) k_adj_fsum(AB A) {
    const PT P = PT::make(A);
    int i, j, k;
    long long p;
    if (!point(A, P, i, j, k, p)) return;
    const long long co = (long long)blockIdx.y * A.cs;
    double l[3], r[3], unused;
#pragma unroll
    for (int d = 0; d < 3; ++d) face_totals<CONS>(A, P, co, i, j, k, d, false, l[d], r[d], unused);
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
template <class PT, bool FS, bool ROLL, bool CONS>
__global__ void __launch_bounds__(NTB) k_adj_trace(AB A) {
    const PT P = PT::make(A);
    int i, j, k;
    long long p;
    if (!point(A, P, i, j, k, p)) return;
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
            face_totals<CONS>(A, P, co, i + (d == 0), j + (d == 1), k + (d == 2), d, false, Ipb,
                        unused0, unused1);
            // face p's totals also give its velocity adjoint for this component
            face_totals<CONS>(A, P, co, i, j, k, d, true, unused2, Imb, ubd[ROLL ? 0 : d]);
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
        if (CONS) s0s += cons_direct(A, P, co, i, j, k);
        A.S0S[co + p] = s0s;
        return;
    }
    double s0t = (r0[0] + r0[1]) + r0[2];
    if (CONS) s0t += cons_direct(A, P, co, i, j, k);
    A.S0S[co + p] = s0t;
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
template <class PT, bool CONS>
__global__ void __launch_bounds__(NTB) k_adj_trace2(AB A) {
    const PT P = PT::make(A);
    int i, j, k;
    long long p;
    if (!point(A, P, i, j, k, p)) return;
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
        face_totals<CONS>(A, P, co, i + (d == 0), j + (d == 1), k + (d == 2), d, false, Ipb, unused0,
                    unused1);
        face_totals<CONS>(A, P, co, i, j, k, d, true, unused2, Imb, ubd);
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
    if (CONS) s0s += cons_direct(A, P, co, i, j, k);
    A.S0S[co + p] = s0s;
}

// The state adjoint of a cell, from the gather g(i, j, k) of the adjoints
// of every use of the state at position (i, j, k).  Periodic: the cell's
// own position.  General layout: a cell on an outflow side also collects
// the positions outside it whose ghost value is a copy of it (foextrap),
// up to the stencils' reach A.fold.
template <class PT, class F>
__device__ __forceinline__ double fold_state(const AB& A, const PT& P, int i, int j, int k, F g) {
    if constexpr (!PT::kGen) {
        return g(i, j, k);
    } else {
        // the range of positions per direction, without runtime-indexed
        // arrays (they would go to scratch memory)
        auto range = [&](int q, int d, int& lo, int& hi) {
            lo = hi = q;
            if (P.per_of(d)) return;
            if (q == 0) lo = -A.fold;
            if (q == P.nn_of(d) - 1) hi = q + A.fold;
        };
        int l0, h0, l1, h1, l2, h2;
        range(i, 0, l0, h0);
        range(j, 1, l1, h1);
        range(k, 2, l2, h2);
        if (l0 == h0 && l1 == h1 && l2 == h2) return g(i, j, k);
        double s = 0.0;
        for (int c = l2; c <= h2; ++c)
            for (int b = l1; b <= h1; ++b)
                for (int a = l0; a <= h0; ++a) s += g(a, b, c);
        return s;
    }
}

// B7-B9 for ppm_type 2: the state adjoint of cell p gathers, per direction,
// the direct limiter terms of the cells p - 2d .. p + 2d and the face
// adjoints of the faces p - d .. p + 2d through their Colella-Sekora values
template <class PT>
__device__ __forceinline__ double state2_at(const AB& A, const PT& P, long long co,
                                            const double* S, int i, int j, int k) {
    double s = A.G[co + P.at(i, j, k)] + A.S0S[co + P.at(i, j, k)];
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
    return s;
}

template <class PT>
__global__ void __launch_bounds__(NTB) k_adj_state2(AB A) {
    const PT P = PT::make(A);
    int i, j, k;
    long long p;
    if (!point(A, P, i, j, k, p) || !P.cell(i, j, k)) return;
    const long long co = (long long)blockIdx.y * A.cs;
    const double* S = A.S + co;
    A.SBF(i, j, k, (int)blockIdx.y) = fold_state(
        A, P, i, j, k, [&](int a, int b, int c) { return state2_at(A, P, co, S, a, b, c); });
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
template <class PT>
__global__ void __launch_bounds__(NTB) k_adj_face(AB A) {
    const PT P = PT::make(A);
    int i, j, k;
    long long p;
    if (!point(A, P, i, j, k, p)) return;
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
template <class PT>
__global__ void __launch_bounds__(NTB) k_adj_slope(AB A) {
    const PT P = PT::make(A);
    int i, j, k;
    long long p;
    if (!point(A, P, i, j, k, p)) return;
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
template <class PT>
__device__ __forceinline__ double state_fused_at(const AB& A, const PT& P, long long co,
                                                 const double* S, int i, int j, int k) {
    double s = A.G[co + P.at(i, j, k)] + A.S0S[co + P.at(i, j, k)];
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
    return s;
}

template <class PT>
__global__ void __launch_bounds__(NTB) k_adj_state_fused(AB A) {
    const PT P = PT::make(A);
    int i, j, k;
    long long p;
    if (!point(A, P, i, j, k, p) || !P.cell(i, j, k)) return;
    const long long co = (long long)blockIdx.y * A.cs;
    const double* S = A.S + co;
    A.SBF(i, j, k, (int)blockIdx.y) = fold_state(
        A, P, i, j, k, [&](int a, int b, int c) { return state_fused_at(A, P, co, S, a, b, c); });
}

// B9: state adjoint
template <class PT>
__global__ void __launch_bounds__(NTB) k_adj_state(AB A) {
    const PT P = PT::make(A);
    int i, j, k;
    long long p;
    if (!point(A, P, i, j, k, p)) return;
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
template <class PT, bool STORED, bool CONS>
__global__ void __launch_bounds__(NTB) k_adj_vel(AB A, DA ub0, DA ub1, DA ub2, int has_n) {
    const PT P = PT::make(A);
    int i, j, k;
    long long p;
    if (!point(A, P, i, j, k, p)) return;
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
                face_totals<CONS>(A, P, co, i, j, k, d, true, slb, srb, ubt);
                ub = ubt + A.ULO[d][co + p];
            }
            s += ub + A.UHI[d][co + cm];
        }
        r[d] = s;
    }
    if constexpr (PT::kGen) {
        // faces 0 .. nn[d] along d (0 .. nn[d]-1 when periodic), cells
        // across it
        const bool ci = i < P.nn0, cj = j < P.nn1, ck = k < P.nn2;
        if (cj && ck && (ci || !P.per0)) ub0(i, j, k) = r[0];
        if (ci && ck && (cj || !P.per1)) ub1(i, j, k) = r[1];
        if (ci && cj && (ck || !P.per2)) ub2(i, j, k) = r[2];
        return;
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

template <class PT, class PC, bool CONS>
int vjp_kernels_t(mx_adv_ctx* c, const AB& A, dim3 g1, dim3 gc, DA ub0, DA ub1, DA ub2,
                  int has_n) {
    const bool ppm2 = c->p.ppm_type == 2;
    // the three-kernel state gather has no ghost folding: the general
    // layout takes the fused one
    const bool state_fused = c->adj_state_fused || PT::kGen;
    if (ppm2)
        MX_LAUNCH(c, "k_tape_ppm", (k_tape_ppm<2, PT>), gc, dim3(NTB), A);
    else
        MX_LAUNCH(c, "k_tape_ppm", (k_tape_ppm<1, PT>), gc, dim3(NTB), A);
    MX_LAUNCH(c, "k_tape_corner", (k_tape_corner<PT, CONS>), gc, dim3(NTB), A);
    MX_LAUNCH(c, "k_tape_edge", (k_tape_edge<PT, CONS>), gc, dim3(NTB), A);
    MX_LAUNCH(c, "k_adj_tterm", (k_adj_tterm<PC, CONS>), gc, dim3(NTB), A);
    if (ppm2) {
        // ppm_type 2: the Colella-Sekora trace and state sweeps (the kernel
        // arrangement options apply to ppm_type 1)
        MX_LAUNCH(c, "k_adj_trace", (k_adj_trace2<PC, CONS>), gc, dim3(NTB), A);
        MX_LAUNCH(c, "k_adj_state", k_adj_state2<PT>, gc, dim3(NTB), A);
    } else if (c->adj_face_sums) {
        MX_LAUNCH(c, "k_adj_fsum", (k_adj_fsum<PT, CONS>), gc, dim3(NTB), A);
        MX_LAUNCH(c, "k_adj_trace", (k_adj_trace<PC, true, false, CONS>), gc, dim3(NTB), A);
    } else if (c->adj_trace_rolled) {
        MX_LAUNCH(c, "k_adj_trace", (k_adj_trace<PC, false, true, CONS>), gc, dim3(NTB), A);
    } else {
        MX_LAUNCH(c, "k_adj_trace", (k_adj_trace<PC, false, false, CONS>), gc, dim3(NTB), A);
    }
    if (ppm2) {
        // the state sweep ran above
    } else if (state_fused) {
        MX_LAUNCH(c, "k_adj_state_fused", k_adj_state_fused<PT>, gc, dim3(NTB), A);
    } else {
        MX_LAUNCH(c, "k_adj_face", k_adj_face<PT>, gc, dim3(NTB), A);
        MX_LAUNCH(c, "k_adj_slope", k_adj_slope<PT>, gc, dim3(NTB), A);
        MX_LAUNCH(c, "k_adj_state", k_adj_state<PT>, gc, dim3(NTB), A);
    }
    if (c->adj_face_sums && !ppm2)
        MX_LAUNCH(c, "k_adj_vel", (k_adj_vel<PT, false, CONS>), g1, dim3(NTB), A, ub0, ub1, ub2,
                  has_n);
    else
        MX_LAUNCH(c, "k_adj_vel", (k_adj_vel<PT, true, CONS>), g1, dim3(NTB), A, ub0, ub1, ub2,
                  has_n);
    return MX_OK;
}

// the tape and reverse kernels of the step adjoint.  PT indexes the tape,
// state and velocity sweeps, PC the transverse-term and trace gathers
// (which run faster with compare indexing on MI355X, measured, whatever n
// is); the velocity adjoint goes to ub0..ub2 (has_n: the periodic FABs
// that hold face n); cons: the conservative form
template <class PT, class PC>
int vjp_kernels(mx_adv_ctx* c, const AB& A, dim3 g1, dim3 gc, DA ub0, DA ub1, DA ub2, int has_n,
                bool cons) {
    return cons ? vjp_kernels_t<PT, PC, true>(c, A, g1, gc, ub0, ub1, ub2, has_n)
                : vjp_kernels_t<PT, PC, false>(c, A, g1, gc, ub0, ub1, ub2, has_n);
}

// arrays of the tape and the reverse sweeps: `take(m)` hands out m
// components of cs points each; returns how many components it took (with
// take == nullptr: counts only)
template <class Take>
long long ab_arrays(const mx_adv_ctx* c, AB& A, int ncomp, bool gen, Take take) {
    const bool ppm2 = c->p.ppm_type == 2;
    const bool fs = c->adj_face_sums && !ppm2;
    const bool staged_state = !ppm2 && !c->adj_state_fused && !gen;
    long long n = 0;
    auto t = [&](double*& dst) {
        dst = take(ncomp);
        n += ncomp;
    };
    for (int d = 0; d < 3; ++d) {
        t(A.IP[d]);
        t(A.IM[d]);
        t(A.SIMH[d]);
        t(A.SE[d]);
        t(A.TB[d]);
        t(A.UBC[d]);
        if (ppm2)
            for (int o = 0; o < 4; ++o) t(A.DV[d][o]);
        if (fs) {
            t(A.SLB[d]);
            t(A.SRB[d]);
            t(A.ULO[d]);
        }
        t(A.FLB[d]);
        t(A.FRB[d]);
        t(A.UHI[d]);
        if (staged_state) {
            t(A.DSA[d]);
            t(A.DSB[d]);
            t(A.DLA[d]);
            t(A.DLB[d]);
            t(A.QM[d]);
            t(A.Q0[d]);
            t(A.QP[d]);
        }
    }
    for (int q = 0; q < 6; ++q) {
        t(A.COR[q]);
        t(A.EB[q]);
    }
    t(A.S0S);
    return n;
}

void ab_factors(const mx_adv_ctx* c, AB& A, int ncomp) {
    A.ncomp = ncomp;
    A.dt = c->p.dt;
    A.rel_eps = c->p.rel_eps;
    const double dt2 = 0.5 * A.dt, dt3 = A.dt / 3.0, dt4 = 0.25 * A.dt, dt6 = A.dt / 6.0;
    A.dt3 = dt3;
    A.dx0 = c->p.dx[0][0];
    for (int d = 0; d < 3; ++d) {
        A.dx[d] = c->p.dx[0][d];
        A.fdt[d] = A.dt / A.dx[d];
        A.f4[d] = dt4 / A.dx[d];
        A.f6[d] = dt6 / A.dx[d];
        A.f2[d] = dt2 / A.dx[d];
        A.f3[d] = dt3 / A.dx[d];
    }
}

// ---------------------------------------------------------------- box lists
// The step over the level-0 box list of a context: the boxes tile the
// domain, each direction periodic or outflow on both sides.
struct StepDomain {
    int lo[3], nn[3], per[3];
};

int box_domain(const mx_adv_ctx* c, int ncomp, StepDomain& D) {
    const mx_adv_params& p = c->p;
    if (p.dim != 3 || (p.ppm_type != 1 && p.ppm_type != 2) || p.ppm_trace_forces != 0)
        return MX_ERR_UNSUPPORTED;
    if (ncomp < 1 || ncomp > MX_MAX_SCAL) return MX_ERR_ARG;
    long long dom = 1;
    for (int d = 0; d < 3; ++d) {
        D.lo[d] = p.domain_lo[0][d];
        D.nn[d] = p.domain_hi[0][d] - p.domain_lo[0][d] + 1;
        if (p.is_periodic[d])
            D.per[d] = 1;
        else if (p.phys_bc_lo[d] == MX_PHYS_OUTFLOW && p.phys_bc_hi[d] == MX_PHYS_OUTFLOW)
            D.per[d] = 0;
        else
            return MX_ERR_UNSUPPORTED;
        if (D.nn[d] < 4) return MX_ERR_UNSUPPORTED;
        dom *= D.nn[d];
    }
    // set_level keeps the boxes disjoint and inside the domain, so they
    // tile it exactly when their cells add up to it
    long long cells = 0;
    for (const mx_box& b : c->grids[0]) cells += npts(hbox(b));
    if (cells != dom) return MX_ERR_UNSUPPORTED;
    return MX_OK;
}

int nfab_ok(const mx_adv_ctx* c, const mx_multifab* m) {
    return m && m->fab && m->nfab == (int)c->grids[0].size();
}

// snew = s + dt (-div(u sedge) + 0) on one box per task
struct UpdTask {
    DA s, snew, u[3], e[3];
    int lo[3], hi[3];
};

__global__ void __launch_bounds__(NTB) k_box_update(const UpdTask* __restrict__ tasks, int ncomp,
                                                    double dt, double dx0, double dx1, double dx2) {
    const UpdTask& t = tasks[blockIdx.y];
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!region_point(idx, t.lo, t.hi, i, j, k)) return;
    for (int c = 0; c < ncomp; ++c) {
        const double fx1 = t.u[0](i + 1, j, k) * t.e[0](i + 1, j, k, c);
        const double fx0 = t.u[0](i, j, k) * t.e[0](i, j, k, c);
        const double fy1 = t.u[1](i, j + 1, k) * t.e[1](i, j + 1, k, c);
        const double fy0 = t.u[1](i, j, k) * t.e[1](i, j, k, c);
        const double fz1 = t.u[2](i, j, k + 1) * t.e[2](i, j, k + 1, c);
        const double fz0 = t.u[2](i, j, k) * t.e[2](i, j, k, c);
        double divterm = (fx1 - fx0) / dx0;
        divterm += (fy1 - fy0) / dx1;
        divterm += (fz1 - fz0) / dx2;
        t.snew(i, j, k, c) = t.s(i, j, k, c) + dt * (-divterm + 0.0);
    }
}

// the dense layout of the general adjoint: E[d] stored points per
// direction, the domain's cell lo at stored index pad[d]
struct Dense {
    double* p;
    long long cs;  // per component
    int E[3], pad[3], dlo[3], nn[3], per[3];
    __device__ __forceinline__ long long at(int x, int y, int z) const {
        return (long long)(x - dlo[0] + pad[0]) +
               (long long)E[0] * ((long long)(y - dlo[1] + pad[1]) +
                                  (long long)E[1] * (long long)(z - dlo[2] + pad[2]));
    }
};

// a box's region of a FAB <-> the dense layout: cells (nodal -1) or the
// faces lo..hi+1 of direction `nodal`; a face position nn[d] along a
// periodic direction is face 0
struct XTask {
    DA f;
    int lo[3], hi[3];
};

template <bool TO_DENSE>
__global__ void __launch_bounds__(NTB) k_box_xfer(const XTask* __restrict__ tasks, Dense D,
                                                  int ncomp) {
    const XTask& t = tasks[blockIdx.y];
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!region_point(idx, t.lo, t.hi, i, j, k)) return;
    int q[3] = {i, j, k};
    for (int d = 0; d < 3; ++d)
        if (D.per[d] && q[d] - D.dlo[d] >= D.nn[d]) q[d] -= D.nn[d];
    const long long o = D.at(q[0], q[1], q[2]);
    for (int c = 0; c < ncomp; ++c) {
        if (TO_DENSE)
            D.p[c * D.cs + o] = t.f(i, j, k, c);
        else
            t.f(i, j, k, c) = D.p[c * D.cs + o];
    }
}

// every stored point of a dense array that is not a valid cell (nodal -1)
// or face of direction `nodal`: the nearest valid cell's value (clamp) or
// zero
__global__ void __launch_bounds__(NTB) k_dense_outside(Dense D, int ncomp, int nodal, int clamp) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long tot = (long long)D.E[0] * D.E[1] * D.E[2];
    if (q >= tot) return;
    const long long r = q / D.E[0];
    const int x[3] = {(int)(q - r * D.E[0]) - D.pad[0], (int)(r % D.E[1]) - D.pad[1],
                      (int)(r / D.E[1]) - D.pad[2]};
    bool inside = true;
    int cl[3];
    for (int d = 0; d < 3; ++d) {
        const int top = D.nn[d] - 1 + (d == nodal && !D.per[d] ? 1 : 0);
        if (x[d] < 0 || x[d] > top) inside = false;
        cl[d] = x[d] < 0 ? 0 : (x[d] > D.nn[d] - 1 ? D.nn[d] - 1 : x[d]);
    }
    if (inside) return;
    const long long from = (long long)(cl[0] + D.pad[0]) +
                           (long long)D.E[0] * ((long long)(cl[1] + D.pad[1]) +
                                                (long long)D.E[1] * (long long)(cl[2] + D.pad[2]));
    for (int c = 0; c < ncomp; ++c) D.p[c * D.cs + q] = clamp ? D.p[c * D.cs + from] : 0.0;
}

// zero the padding shell (the stored points outside the visited ones) of
// every array of a table, one array per blockIdx.y
__global__ void __launch_bounds__(NTB) k_shell_zero(double* const* __restrict__ arrs, int E0,
                                                    int E1, int E2, int x0, int x1, int x2,
                                                    int p0, int p1, int p2) {
    double* a = arrs[blockIdx.y];
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long plane = (long long)E0 * E1;
    const long long n1 = 2LL * p2 * plane;              // low and high z slabs
    const long long n2 = (long long)x2 * 2 * p1 * E0;   // y slabs inside them
    const long long n3 = (long long)x2 * x1 * 2 * p0;   // x slabs inside both
    int i, j, k;
    if (q < n1) {
        const int t = (int)(q / plane);
        const long long r = q - (long long)t * plane;
        k = t < p2 ? t : x2 + t;
        j = (int)(r / E0);
        i = (int)(r - (long long)j * E0);
    } else if ((q -= n1) < n2) {
        const long long row = 2LL * p1 * E0;
        const int t = (int)(q / row);
        const long long r = q - (long long)t * row;
        const int rr = (int)(r / E0);
        k = p2 + t;
        j = rr < p1 ? rr : x1 + rr;
        i = (int)(r - (long long)rr * E0);
    } else if ((q -= n2) < n3) {
        const int t = (int)(q / (2 * p0));
        const int rr = (int)(q - (long long)t * 2 * p0);
        k = p2 + t / x1;
        j = p1 + t % x1;
        i = rr < p0 ? rr : x0 + rr;
    } else {
        return;
    }
    a[(long long)i + (long long)E0 * ((long long)j + (long long)E1 * k)] = 0.0;
}

dim3 grid1(long long npts) { return dim3((unsigned)((npts + NTB - 1) / NTB)); }

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

// did the views taken since ws_reserve stay inside the workspace?
bool ws_fits(const mx_adv_ctx* c) {
    if (c->dry) return true;
    return c->ws_off <= (c->ext_ws ? c->ext_ws_cap : c->ws_cap);
}

mx_fab dense_desc(double* p, const int ext[3], const int begin[3], int ncomp) {
    mx_fab f;
    f.p = p;
    f.jstride = ext[0];
    f.kstride = (long long)ext[0] * ext[1];
    f.nstride = (long long)ext[0] * ext[1] * ext[2];
    for (int q = 0; q < 3; ++q) {
        f.begin[q] = begin[q];
        f.end[q] = begin[q] + ext[q];
    }
    f.ncomp = ncomp;
    return f;
}

}  // namespace

extern "C" {

int mx_adv_scal_step(mx_adv_ctx* c, mx_fab* s, mx_fab* umac, mx_fab* snew, int ncomp,
                     int is_conservative) {
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
        const int ex[3] = {n + (d == 0), n + (d == 1), n + (d == 2)}, b0[3] = {0, 0, 0};
        se[d] = dense_desc(ws_take(c, (size_t)(fn * ncomp)), ex, b0, ncomp);
    }
    const int exf[3] = {n + 2, n + 2, n + 2}, bf[3] = {-1, -1, -1};
    fo = dense_desc(ws_take(c, (size_t)(cn * ncomp)), exf, bf, ncomp);
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
    rc = make_edge_scal(c, 0, &sm, &sem, &um, &fm, 0, bcs.data(), ncomp, 0, 0, ncomp,
                        is_conservative != 0);
    c->fused_edge = saved;
    if (rc) return rc;
    MX_LAUNCH(c, "k_step_update", k_step_update, grid1((long long)n * n * n), dim3(NTB), to_da(*s),
              to_da(*snew), to_da(umac[0]), to_da(umac[1]), to_da(umac[2]), to_da(se[0]),
              to_da(se[1]), to_da(se[2]), n, ncomp, c->p.dt, c->p.dx[0][0], c->p.dx[0][1],
              c->p.dx[0][2]);
    return MX_OK;
}

int mx_adv_scal_step_vjp(mx_adv_ctx* c, mx_fab* s, mx_fab* umac, const mx_fab* gbar,
                         mx_fab* sbar, mx_fab* ubar, int ncomp, int is_conservative) {
    if (!c || !s || !umac || !gbar || !sbar || !ubar) return MX_ERR_ARG;
    int n = 0;
    int rc = step_check(c, s, umac, ncomp, n, false);
    if (rc) return rc;
    HBox v = hbox(c->grids[0][0]);
    if (!covers(*gbar, v, ncomp) || !covers(*sbar, v, ncomp)) return MX_ERR_BOX;
    for (int d = 0; d < 3; ++d)
        if (!covers(ubar[d], v, 1)) return MX_ERR_BOX;
    const long long n3 = (long long)n * n * n;
    AB A;
    std::memset(&A, 0, sizeof(A));
    // dense arrays: S and G (ncomp each), the intermediates and adjoints, U (3)
    const size_t per = (size_t)n3 * sizeof(double) + 256;
    const long long count = ab_arrays(c, A, ncomp, false, [](int) { return (double*)nullptr; });
    if ((rc = ws_reserve(c, per * (size_t)(count + 2 * ncomp + 3) + 4096))) return rc;
    auto take = [&](int m) { return ws_take(c, (size_t)n3 * m); };
    // inputs that are already dense (ncomp, n, n, n) arrays are read in place
    const bool sd = dense_fab(*s, n, ncomp), gd = dense_fab(*gbar, n, ncomp);
    bool ud[3];
    for (int d = 0; d < 3; ++d) ud[d] = dense_fab(umac[d], n, 1);
    double* S = sd ? s->p : take(ncomp);
    double* G = gd ? gbar->p : take(ncomp);
    for (int d = 0; d < 3; ++d) A.U[d] = ud[d] ? umac[d].p : take(1);
    ab_arrays(c, A, ncomp, false, take);
    A.S = S;
    A.G = G;
    A.SBF = to_da(*sbar);
    A.cs = n3;
    A.npts = n3;
    A.n = n;
    for (int d = 0; d < 3; ++d) A.ext[d] = n;
    ab_factors(c, A, ncomp);

    A.xcd = c->adj_xcd;
    dim3 g1 = grid1(n3);
    if (A.xcd) g1.x = (g1.x + 7u) & ~7u;
    const dim3 gc(g1.x, (unsigned)ncomp);
    if (!sd) MX_LAUNCH(c, "k_to_dense", k_to_dense, g1, dim3(NTB), to_da(*s), 0, ncomp, n, S);
    if (!gd)
        MX_LAUNCH(c, "k_to_dense", k_to_dense, g1, dim3(NTB), to_da(*gbar), 0, ncomp, n, G);
    for (int d = 0; d < 3; ++d)
        if (!ud[d])
            MX_LAUNCH(c, "k_to_dense", k_to_dense, g1, dim3(NTB), to_da(umac[d]), 0, 1, n,
                      const_cast<double*>(A.U[d]));
    int has_n = 0;
    for (int d = 0; d < 3; ++d)
        if (ubar[d].end[d] > n) has_n |= 1 << d;
    const DA u0 = to_da(ubar[0]), u1 = to_da(ubar[1]), u2 = to_da(ubar[2]);
    const bool p2 = c->adj_pow2 && (n & (n - 1)) == 0;
    const bool cons = is_conservative != 0;
    return p2 ? vjp_kernels<PerT<true>, PerT<false>>(c, A, g1, gc, u0, u1, u2, has_n, cons)
              : vjp_kernels<PerT<false>, PerT<false>>(c, A, g1, gc, u0, u1, u2, has_n, cons);
}

int mx_adv_scal_step_boxes(mx_adv_ctx* c, mx_multifab* s, mx_face_multifab* umac,
                           mx_multifab* snew, int ncomp, int is_conservative) {
    if (!c || !s || !umac || !snew) return MX_ERR_ARG;
    StepDomain D;
    int rc = box_domain(c, ncomp, D);
    if (rc) return rc;
    const auto& grids = c->grids[0];
    const int nb = (int)grids.size();
    if (!nfab_ok(c, s) || !nfab_ok(c, snew)) return MX_ERR_ARG;
    for (int d = 0; d < 3; ++d)
        if (!nfab_ok(c, &umac->dir[d])) return MX_ERR_ARG;
    const int g = c->p.ppm_type == 2 ? 4 : 3;
    for (int b = 0; b < nb; ++b) {
        const HBox v = hbox(grids[b]);
        if (!covers(s->fab[b], grow(v, 3, g), ncomp) || !covers(snew->fab[b], v, ncomp))
            return MX_ERR_BOX;
        for (int d = 0; d < 3; ++d)
            if (!covers(umac->dir[d].fab[b], grow(nodal(v, d), 3, 1), 1)) return MX_ERR_BOX;
    }
    // ghost cells and faces: FillPatch (periodic images, foextrap beyond an
    // outflow side) and FillBoundary + FillUmacGhost
    std::vector<mx_bcrec> bcs(ncomp);
    for (auto& b : bcs)
        for (int d = 0; d < 3; ++d) b.lo[d] = b.hi[d] = D.per[d] ? MX_BC_INT_DIR : MX_BC_FOEXTRAP;
    if ((rc = mx_adv_fill_patch(c, 0, s, s, 0, 0, ncomp, bcs.data(), 0))) return rc;
    if ((rc = fill_face_ghosts(c, 0, umac))) return rc;
    // edge states per box (no ghosts) and a zero force with one ghost
    size_t need = 8192;
    for (int b = 0; b < nb; ++b) {
        const HBox v = hbox(grids[b]);
        for (int d = 0; d < 3; ++d) need += (size_t)npts(nodal(v, d)) * ncomp * 8 + 256;
        need += (size_t)npts(grow(v, 3, 1)) * ncomp * 8 + 256;
    }
    if ((rc = ws_reserve(c, need))) return rc;
    std::vector<mx_fab> se[3], fo(nb);
    long long maxpts = 0;
    for (int b = 0; b < nb; ++b) {
        const HBox v = hbox(grids[b]);
        for (int d = 0; d < 3; ++d) {
            const HBox f = nodal(v, d);
            const int ex[3] = {f.hi[0] - f.lo[0] + 1, f.hi[1] - f.lo[1] + 1,
                               f.hi[2] - f.lo[2] + 1};
            se[d].push_back(dense_desc(ws_take(c, (size_t)npts(f) * ncomp), ex, f.lo, ncomp));
        }
        maxpts = std::max(maxpts, npts(v));
    }
    // the zero force of every box in one span (one fill)
    for (int b = 0; b < nb; ++b) {
        const HBox gv = grow(hbox(grids[b]), 3, 1);
        const int ex[3] = {gv.hi[0] - gv.lo[0] + 1, gv.hi[1] - gv.lo[1] + 1,
                           gv.hi[2] - gv.lo[2] + 1};
        fo[b] = dense_desc(ws_take(c, (size_t)npts(gv) * ncomp), ex, gv.lo, ncomp);
    }
    if (!ws_fits(c)) return MX_ERR_NOMEM;
    const long long span = (long long)(fo[nb - 1].p + fo[nb - 1].nstride * ncomp - fo[0].p);
    MX_LAUNCH(c, "k_fill_zero", k_fill_zero, grid1(span), dim3(NTB), fo[0].p, span);
    mx_multifab fm{nb, fo.data()};
    mx_face_multifab sem;
    for (int d = 0; d < 3; ++d) sem.dir[d] = {nb, se[d].data()};
    const int saved = c->fused_edge;
    c->fused_edge = 1;
    rc = make_edge_scal(c, 0, s, &sem, umac, &fm, 0, bcs.data(), ncomp, 0, 0, ncomp,
                        is_conservative != 0);
    c->fused_edge = saved;
    if (rc) return rc;
    std::vector<UpdTask> tasks(nb);
    for (int b = 0; b < nb; ++b) {
        const HBox v = hbox(grids[b]);
        UpdTask& t = tasks[b];
        t.s = to_da(s->fab[b]);
        t.snew = to_da(snew->fab[b]);
        for (int d = 0; d < 3; ++d) {
            t.u[d] = to_da(umac->dir[d].fab[b]);
            t.e[d] = to_da(se[d][b]);
            t.lo[d] = v.lo[d];
            t.hi[d] = v.hi[d];
        }
    }
    Table tab(c);
    const size_t off = tab.add(tasks);
    if ((rc = tab.upload())) return rc;
    MX_LAUNCH(c, "k_box_update", k_box_update, grid_for(maxpts, nb, 1, NTB), dim3(NTB),
              tab.dev<UpdTask>(off), ncomp, c->p.dt, c->p.dx[0][0], c->p.dx[0][1],
              c->p.dx[0][2]);
    return MX_OK;
}

int mx_adv_scal_step_vjp_boxes(mx_adv_ctx* c, const mx_multifab* s, const mx_face_multifab* umac,
                               const mx_multifab* gbar, mx_multifab* sbar,
                               mx_face_multifab* ubar, int ncomp, int is_conservative) {
    if (!c || !s || !umac || !gbar || !sbar || !ubar) return MX_ERR_ARG;
    StepDomain D;
    int rc = box_domain(c, ncomp, D);
    if (rc) return rc;
    const auto& grids = c->grids[0];
    const int nb = (int)grids.size();
    if (!nfab_ok(c, s) || !nfab_ok(c, gbar) || !nfab_ok(c, sbar)) return MX_ERR_ARG;
    for (int d = 0; d < 3; ++d)
        if (!nfab_ok(c, &umac->dir[d]) || !nfab_ok(c, &ubar->dir[d])) return MX_ERR_ARG;
    long long maxpts = 0;
    for (int b = 0; b < nb; ++b) {
        const HBox v = hbox(grids[b]);
        if (!covers(s->fab[b], v, ncomp) || !covers(gbar->fab[b], v, ncomp) ||
            !covers(sbar->fab[b], v, ncomp))
            return MX_ERR_BOX;
        for (int d = 0; d < 3; ++d)
            if (!covers(umac->dir[d].fab[b], nodal(v, d), 1) ||
                !covers(ubar->dir[d].fab[b], nodal(v, d), 1))
                return MX_ERR_BOX;
        maxpts = std::max(maxpts, npts(nodal(nodal(nodal(v, 0), 1), 2)));
    }
    // the layout: along an outflow direction positions 0..nn (face nn) and
    // the reach of the stencils (6 points: ghost positions up to 3 outside,
    // each read 3 further) on both sides
    constexpr int PAD = 6;
    AB A;
    std::memset(&A, 0, sizeof(A));
    Dense L;
    L.p = nullptr;
    long long cs = 1, npt = 1;
    for (int d = 0; d < 3; ++d) {
        L.pad[d] = D.per[d] ? 0 : PAD;
        A.ext[d] = D.nn[d] + (D.per[d] ? 0 : 1);
        L.E[d] = A.ext[d] + 2 * L.pad[d];
        L.dlo[d] = D.lo[d];
        L.nn[d] = D.nn[d];
        L.per[d] = D.per[d];
        cs *= L.E[d];
        npt *= A.ext[d];
        (d == 0 ? A.gen.nn0 : (d == 1 ? A.gen.nn1 : A.gen.nn2)) = D.nn[d];
        (d == 0 ? A.gen.per0 : (d == 1 ? A.gen.per1 : A.gen.per2)) = D.per[d];
        (d == 0 ? A.gen.pad0 : (d == 1 ? A.gen.pad1 : A.gen.pad2)) = L.pad[d];
    }
    if (cs > 0x7fffffffLL || npt > 0xffffffffLL) return MX_ERR_UNSUPPORTED;
    L.cs = cs;
    A.gen.sj = L.E[0];
    A.gen.sk = L.E[0] * L.E[1];
    A.cs = cs;
    A.npts = npt;
    A.fold = 3;
    A.n = D.nn[0];
    const long long ncell = (long long)D.nn[0] * D.nn[1] * D.nn[2];
    long long nface[3];
    for (int d = 0; d < 3; ++d)
        nface[d] = ncell / D.nn[d] * (D.nn[d] + (D.per[d] ? 0 : 1));
    const long long count = ab_arrays(c, A, ncomp, true, [](int) { return (double*)nullptr; });
    size_t need = (size_t)(2 * ncomp + 3) * ((size_t)cs * 8 + 256) +
                  (size_t)ncell * ncomp * 8 + (nface[0] + nface[1] + nface[2]) * 8 + 8192;
    if ((rc = ws_reserve(c, need))) return rc;
    // the intermediates live in the context's adjoint buffer, whose padding
    // is zeroed when the buffer or the layout changes
    const size_t abytes = (size_t)count * (size_t)cs * 8 + 256 * (size_t)(count + 1);
    const long long layout[8] = {L.E[0], L.E[1], L.E[2], A.ext[0], A.ext[1], A.ext[2], count,
                                 ncomp};
    bool rezero = std::memcmp(layout, c->adj_layout, sizeof(layout)) != 0;
    if (abytes > c->adj_cap) {
        MX_HIP_CHECK(hipStreamSynchronize(c->stream));
        if (c->adj_buf) MX_HIP_CHECK(hipFree(c->adj_buf));
        c->adj_buf = nullptr;
        c->adj_cap = 0;
        if (hipMalloc((void**)&c->adj_buf, abytes) != hipSuccess) {
            c->adj_buf = nullptr;
            return MX_ERR_NOMEM;
        }
        c->adj_cap = abytes;
        rezero = true;
    }
    std::vector<double*> padded;  // every array whose shell the sweeps read as zero
    size_t aoff = 0;
    auto take = [&](int m) {
        aoff = (aoff + 255) & ~(size_t)255;
        double* p = reinterpret_cast<double*>(c->adj_buf + aoff);
        aoff += (size_t)cs * m * sizeof(double);
        for (int q = 0; q < m; ++q) padded.push_back(p + (size_t)q * cs);
        return p;
    };
    double* S = ws_take(c, (size_t)cs * ncomp);
    double* G = ws_take(c, (size_t)cs * ncomp);
    double* U[3];
    for (int d = 0; d < 3; ++d) U[d] = ws_take(c, (size_t)cs);
    ab_arrays(c, A, ncomp, true, take);
    // outputs on the domain's cells and faces
    const int zero3[3] = {0, 0, 0};
    mx_fab sbd = dense_desc(ws_take(c, (size_t)ncell * ncomp), D.nn, zero3, ncomp);
    mx_fab ubd[3];
    for (int d = 0; d < 3; ++d) {
        int ex[3] = {D.nn[0], D.nn[1], D.nn[2]};
        ex[d] += D.per[d] ? 0 : 1;
        ubd[d] = dense_desc(ws_take(c, (size_t)nface[d]), ex, zero3, 1);
    }
    if (!ws_fits(c) || aoff > c->adj_cap) return MX_ERR_NOMEM;
    A.S = S;
    A.G = G;
    for (int d = 0; d < 3; ++d) A.U[d] = U[d];
    A.SBF = to_da(sbd);
    ab_factors(c, A, ncomp);

    // gather the inputs: cells of s and gbar, faces lo..hi (and the high
    // domain face) of umac
    std::vector<XTask> cells_s(nb), cells_g(nb), faces[3];
    for (int b = 0; b < nb; ++b) {
        const HBox v = hbox(grids[b]);
        cells_s[b].f = to_da(s->fab[b]);
        cells_g[b].f = to_da(gbar->fab[b]);
        for (int q = 0; q < 3; ++q) {
            cells_s[b].lo[q] = cells_g[b].lo[q] = v.lo[q];
            cells_s[b].hi[q] = cells_g[b].hi[q] = v.hi[q];
        }
        for (int d = 0; d < 3; ++d) {
            XTask t;
            t.f = to_da(umac->dir[d].fab[b]);
            for (int q = 0; q < 3; ++q) {
                t.lo[q] = v.lo[q];
                t.hi[q] = v.hi[q];
            }
            if (!D.per[d] && v.hi[d] == D.lo[d] + D.nn[d] - 1) t.hi[d] += 1;
            faces[d].push_back(t);
        }
    }
    Table tab(c);
    const size_t o_s = tab.add(cells_s), o_g = tab.add(cells_g);
    size_t o_f[3];
    for (int d = 0; d < 3; ++d) o_f[d] = tab.add(faces[d]);
    const size_t o_p = tab.add(padded);
    if ((rc = tab.upload())) return rc;
    const dim3 gx = grid_for(maxpts, nb, 1, NTB);
    auto dense_of = [&](double* p) {
        Dense q = L;
        q.p = p;
        return q;
    };
    MX_LAUNCH(c, "k_box_xfer", k_box_xfer<true>, gx, dim3(NTB), tab.dev<XTask>(o_s), dense_of(S),
              ncomp);
    MX_LAUNCH(c, "k_box_xfer", k_box_xfer<true>, gx, dim3(NTB), tab.dev<XTask>(o_g), dense_of(G),
              ncomp);
    for (int d = 0; d < 3; ++d)
        MX_LAUNCH(c, "k_box_xfer", k_box_xfer<true>, gx, dim3(NTB), tab.dev<XTask>(o_f[d]),
                  dense_of(U[d]), 1);
    const dim3 ga = grid1(cs);
    MX_LAUNCH(c, "k_dense_outside", k_dense_outside, ga, dim3(NTB), dense_of(S), ncomp, -1, 1);
    MX_LAUNCH(c, "k_dense_outside", k_dense_outside, ga, dim3(NTB), dense_of(G), ncomp, -1, 0);
    for (int d = 0; d < 3; ++d)
        MX_LAUNCH(c, "k_dense_outside", k_dense_outside, ga, dim3(NTB), dense_of(U[d]), 1, d, 0);
    const long long shell = cs - npt;
    if (shell > 0 && rezero)
        MX_LAUNCH(c, "k_shell_zero", k_shell_zero, dim3(grid1(shell).x, (unsigned)padded.size()),
                  dim3(NTB), tab.dev<double*>(o_p), L.E[0], L.E[1], L.E[2], A.ext[0], A.ext[1],
                  A.ext[2], L.pad[0], L.pad[1], L.pad[2]);
    std::memcpy(c->adj_layout, layout, sizeof(layout));

    A.xcd = c->adj_xcd;
    dim3 g1 = grid1(npt);
    if (A.xcd) g1.x = (g1.x + 7u) & ~7u;
    const dim3 gc(g1.x, (unsigned)ncomp);
    if ((rc = vjp_kernels<GenT, GenT>(c, A, g1, gc, to_da(ubd[0]), to_da(ubd[1]), to_da(ubd[2]),
                                      0, is_conservative != 0)))
        return rc;

    // scatter: sbar on every box's cells, ubar on every face a box holds
    // (faces shared by two boxes, and periodic images, get the same value)
    std::vector<XTask> out_s(nb), out_u[3];
    for (int b = 0; b < nb; ++b) {
        const HBox v = hbox(grids[b]);
        out_s[b].f = to_da(sbar->fab[b]);
        for (int q = 0; q < 3; ++q) {
            out_s[b].lo[q] = v.lo[q];
            out_s[b].hi[q] = v.hi[q];
        }
        for (int d = 0; d < 3; ++d) {
            XTask t;
            t.f = to_da(ubar->dir[d].fab[b]);
            const HBox f = nodal(v, d);
            for (int q = 0; q < 3; ++q) {
                t.lo[q] = f.lo[q];
                t.hi[q] = f.hi[q];
            }
            out_u[d].push_back(t);
        }
    }
    Table tab2(c);
    const size_t q_s = tab2.add(out_s);
    size_t q_u[3];
    for (int d = 0; d < 3; ++d) q_u[d] = tab2.add(out_u[d]);
    if ((rc = tab2.upload())) return rc;
    Dense LS = L, LU[3];
    LS.p = sbd.p;
    LS.cs = ncell;
    for (int q = 0; q < 3; ++q) {
        LS.E[q] = D.nn[q];
        LS.pad[q] = 0;
    }
    MX_LAUNCH(c, "k_box_xfer", k_box_xfer<false>, gx, dim3(NTB), tab2.dev<XTask>(q_s), LS, ncomp);
    for (int d = 0; d < 3; ++d) {
        LU[d] = LS;
        LU[d].p = ubd[d].p;
        LU[d].cs = nface[d];
        LU[d].E[d] = D.nn[d] + (D.per[d] ? 0 : 1);
        MX_LAUNCH(c, "k_box_xfer", k_box_xfer<false>, gx, dim3(NTB), tab2.dev<XTask>(q_u[d]),
                  LU[d], 1);
    }
    return MX_OK;
}

}  // extern "C"
