// oracle_step.cpp -- CPU oracle of the differentiable scalar advection step
// and its forward-mode derivative (test infrastructure only, see
// oracle_core.H).
//
// The step, on the cells of a domain with a fixed MAC velocity, each
// direction periodic or outflow (foextrap ghost cells and the foextrap
// rules at the domain faces), for each scalar component:
//   sedge = MakeEdgeScal(s, umac, force = 0, advective form, ppm_type 1 or 2)
//           (MaestroMakeEdgeScalars.cpp: PPM, Predictor :662-863, Transverse
//            :865-1359, Edges :1361-1667; interior PPM formulas: no
//            ext_dir or hoextrap sides)
//   F     = umac * sedge                     (MakeRhoHFlux, predict_rhohprime
//                                             with rhoh0_edge = 0,
//                                             MaestroMakeFlux.cpp:699-708)
//   snew  = s + dt * (-div F + 0)            (UpdateScal, MaestroAdvection.cpp:120-135)
// (the enthalpy step of MaestroEnthalpyAdvance.cpp:143-147; the same
// arithmetic as a species step with species_pred_type 2 and an advective
// prediction, which the oracle's stages are compared against)
// written once for a number type T: double gives the step itself, Dual (a
// value and one directional derivative) gives the Jacobian-vector product
// with the limiter branches taken at the primal values.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "oracle_core.H"

namespace mxo {

namespace {

struct Dual {
    double v, d;
};
inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dual operator/(Dual a, Dual b) {
    return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)};
}
inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
inline Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
inline Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
inline Dual operator+(Dual a, double b) { return {a.v + b, a.d}; }
inline Dual operator+(double a, Dual b) { return {a + b.v, b.d}; }
inline Dual operator-(double a, Dual b) { return {a - b.v, -b.d}; }
inline Dual operator-(Dual a, double b) { return {a.v - b, a.d}; }
inline Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }

// Op-counting number (for the algorithmic FP64 operation count of the step):
// every +, -, *, / on it, every comparison of its value, |x| and copysign
// counts one operation; negation counts none.  val() of it is a proxy whose
// comparisons count one each.
struct Cnt {
    double v;
};
long long g_ops = 0;
struct CmpD {
    double v;
    operator double() const { return v; }
};
inline bool operator<(CmpD a, CmpD b) { ++g_ops; return a.v < b.v; }
inline bool operator<(CmpD a, double b) { ++g_ops; return a.v < b; }
inline bool operator>(CmpD a, double b) { ++g_ops; return a.v > b; }
inline bool operator>(CmpD a, CmpD b) { ++g_ops; return a.v > b.v; }
inline bool operator>=(CmpD a, double b) { ++g_ops; return a.v >= b; }
inline bool operator<=(CmpD a, double b) { ++g_ops; return a.v <= b; }
inline bool operator>=(CmpD a, CmpD b) { ++g_ops; return a.v >= b.v; }
inline Cnt operator+(Cnt a, Cnt b) { ++g_ops; return {a.v + b.v}; }
inline Cnt operator-(Cnt a, Cnt b) { ++g_ops; return {a.v - b.v}; }
inline Cnt operator*(Cnt a, Cnt b) { ++g_ops; return {a.v * b.v}; }
inline Cnt operator/(Cnt a, Cnt b) { ++g_ops; return {a.v / b.v}; }
inline Cnt operator-(Cnt a) { return {-a.v}; }
inline Cnt operator*(double a, Cnt b) { ++g_ops; return {a * b.v}; }
inline Cnt operator*(Cnt a, double b) { ++g_ops; return {a.v * b}; }
inline Cnt operator+(Cnt a, double b) { ++g_ops; return {a.v + b}; }
inline Cnt operator+(double a, Cnt b) { ++g_ops; return {a + b.v}; }
inline Cnt operator-(double a, Cnt b) { ++g_ops; return {a - b.v}; }
inline Cnt operator-(Cnt a, double b) { ++g_ops; return {a.v - b}; }
inline Cnt operator/(Cnt a, double b) { ++g_ops; return {a.v / b}; }

inline double val(double x) { return x; }
inline double val(Dual x) { return x.v; }
inline CmpD val(Cnt x) { return {x.v}; }
inline double mk(double x, double) { return x; }
template <class T>
inline T cst(double x);
template <>
inline double cst<double>(double x) { return x; }
template <>
inline Dual cst<Dual>(double x) { return {x, 0.0}; }
template <>
inline Cnt cst<Cnt>(double x) { return {x}; }

// amrex::min / max (the first argument wins ties), |x|, copysign(1, x)
template <class T>
inline T tmin(T a, T b) { return (val(b) < val(a)) ? b : a; }
template <class T>
inline T tmax(T a, T b) { return (val(a) < val(b)) ? b : a; }
inline double tabs(double a) { return std::fabs(a); }
inline Dual tabs(Dual a) { return std::signbit(a.v) ? Dual{-a.v, -a.d} : a; }
inline Cnt tabs(Cnt a) {
    ++g_ops;
    return {std::fabs(a.v)};
}
template <class T>
inline double sgn1(T a) { return std::copysign(1.0, val(a)); }
inline double sgn1(Cnt a) {
    ++g_ops;
    return std::copysign(1.0, a.v);
}

template <class T>
T vl(T a, T b, T c) {  // van Leer slope (MaestroPPM.cpp:46-57)
    T dsc = 0.5 * (c - a);
    T dsl = 2.0 * (b - a);
    T dsr = 2.0 * (c - b);
    T r = cst<T>(0.0);
    if (val(dsl * dsr) > 0.0) r = sgn1(dsc) * tmin(tabs(dsc), tmin(tabs(dsl), tabs(dsr)));
    return r;
}

// PPM type 1 sm, sp of the cell with stencil v[-2..2] (interior formulas)
template <class T>
void ppm1(const T* v, T& sm, T& sp) {
    T dl = vl(v[-2], v[-1], v[0]);
    T dc = vl(v[-1], v[0], v[1]);
    T dr = vl(v[0], v[1], v[2]);
    sm = 0.5 * (v[0] + v[-1]) - (dc - dl) / 6.0;
    sm = tmax(sm, tmin(v[0], v[-1]));
    sm = tmin(sm, tmax(v[0], v[-1]));
    sp = 0.5 * (v[1] + v[0]) - (dr - dc) / 6.0;
    sp = tmax(sp, tmin(v[1], v[0]));
    sp = tmin(sp, tmax(v[1], v[0]));
    const T s0 = v[0];
    if (val((sp - s0) * (s0 - sm)) <= 0.0) {
        sp = s0;
        sm = s0;
    } else if (val(tabs(sp - s0)) >= val(2.0 * tabs(sm - s0))) {
        sp = 3.0 * s0 - 2.0 * sm;
    } else if (val(tabs(sm - s0)) >= val(2.0 * tabs(sp - s0))) {
        sm = 3.0 * s0 - 2.0 * sp;
    }
}

inline double tsqrt(double a) { return std::sqrt(a); }
inline Dual tsqrt(Dual a) {
    const double r = std::sqrt(a.v);
    return {r, r > 0.0 ? a.d / (2.0 * r) : 0.0};
}
inline Cnt tsqrt(Cnt a) {
    ++g_ops;
    return {std::sqrt(a.v)};
}

// PPM type 2 (Colella-Sekora) face value between cells sb and sc, from the
// cells sa, sb, sc, sd (MaestroPPM.cpp:356-381, interior formula)
template <class T>
T cs_face(T sa, T sb, T sc, T sd) {
    const double C = 1.25;
    T e = (7.0 / 12.0) * (sb + sc) - (1.0 / 12.0) * (sa + sd);
    if (val((e - sb) * (sc - e)) < 0.0) {
        T D2 = 3.0 * (sb - 2.0 * e + sc);
        T D2L = sa - 2.0 * sb + sc;
        T D2R = sb - 2.0 * sc + sd;
        double sgn = sgn1(D2);
        T D2LIM = sgn * tmax(tmin(C * sgn * D2L, tmin(C * sgn * D2R, sgn * D2)), cst<T>(0.0));
        e = 0.5 * (sb + sc) - D2LIM / 6.0;
    }
    return e;
}

// PPM type 2 cell limiter (MaestroPPM.cpp:383-441, interior: threshold 1e-10);
// v[-2..2] = s(cell-2 .. cell+2), faces fl, f0, fr, frr = cell-1 .. cell+2
template <class T>
void cs_cell(const T* v, T fl, T f0, T fr, T frr, T& sm, T& sp) {
    const double C = 1.25;
    const T s0 = v[0];
    T alphap = fr - s0;
    T alpham = f0 - s0;
    const bool bigp = val(tabs(alphap)) > val(2.0 * tabs(alpham));
    const bool bigm = val(tabs(alpham)) > val(2.0 * tabs(alphap));
    bool extremum = false;
    if (val(alpham * alphap) >= 0.0) {
        extremum = true;
    } else if (bigp || bigm) {
        T dafacem = f0 - fl;
        T dafacep = frr - fr;
        T dabarm = s0 - v[-1];
        T dabarp = v[1] - s0;
        T dafacemin = tmin(tabs(dafacem), tabs(dafacep));
        T dabarmin = tmin(tabs(dabarm), tabs(dabarp));
        T dachkm, dachkp;
        if (val(dafacemin) >= val(dabarmin)) {
            dachkm = dafacem;
            dachkp = dafacep;
        } else {
            dachkm = dabarm;
            dachkp = dabarp;
        }
        extremum = val(dachkm * dachkp) <= 0.0;
    }
    if (extremum) {
        T D2 = 6.0 * (alpham + alphap);
        T D2L = v[-2] - 2.0 * v[-1] + s0;
        T D2R = s0 - 2.0 * v[1] + v[2];
        T D2C = v[-1] - 2.0 * s0 + v[1];
        double sgn = sgn1(D2);
        T D2LIM = tmax(tmin(sgn * D2, tmin(C * sgn * D2L, tmin(C * sgn * D2R, C * sgn * D2C))),
                       cst<T>(0.0));
        T D2ABS = tmax(tabs(D2), cst<T>(1.e-10));
        alpham = alpham * D2LIM / D2ABS;
        alphap = alphap * D2LIM / D2ABS;
    } else {
        if (bigp) {
            double sgn = sgn1(alpham);
            T amax = -alphap * alphap / (4.0 * (alpham + alphap));
            T delam = v[-1] - s0;
            if (val(sgn * amax) >= val(sgn * delam)) {
                if (val(sgn * (delam - alpham)) >= 1.e-10) {
                    alphap = -2.0 * delam - 2.0 * sgn * tsqrt(delam * delam - delam * alpham);
                } else {
                    alphap = -2.0 * alpham;
                }
            }
        }
        if (bigm) {
            double sgn = sgn1(alphap);
            T amax = -alpham * alpham / (4.0 * (alpham + alphap));
            T delap = v[1] - s0;
            if (val(sgn * amax) >= val(sgn * delap)) {
                if (val(sgn * (delap - alphap)) >= 1.e-10) {
                    alpham = -2.0 * delap - 2.0 * sgn * tsqrt(delap * delap - delap * alphap);
                } else {
                    alpham = -2.0 * alphap;
                }
            }
        }
    }
    sm = s0 + alpham;
    sp = s0 + alphap;
}

template <class T>
void trace(T s0, T sm, T sp, T uhi, T ulo, double dt, double h, double rel_eps, T& Ip, T& Im) {
    T s6 = 6.0 * s0 - 3.0 * (sm + sp);
    T sigma = tabs(uhi) * dt / h;
    if (val(uhi) > rel_eps)
        Ip = sp - 0.5 * sigma * (sp - sm - (1.0 - 2.0 / 3.0 * sigma) * s6);
    else
        Ip = s0;
    sigma = tabs(ulo) * dt / h;
    if (val(ulo) < -rel_eps)
        Im = sm + 0.5 * sigma * (sp - sm + (1.0 - 2.0 / 3.0 * sigma) * s6);
    else
        Im = s0;
}

template <class T>
T upwind(T l, T r, T u, double thr) {
    T res = (val(u) > 0.0) ? l : r;
    res = (val(tabs(u)) > thr) ? res : 0.5 * (l + r);
    return res;
}

// The step's index space: n[d] cells per direction, each direction periodic
// (faces n[d] and 0 coincide, cells wrap) or outflow (foextrap: faces
// 0 .. n[d], a ghost cell takes the value of the nearest valid cell).
// Cell arrays hold the n0 n1 n2 valid cells, the face arrays of direction d
// one more point along d when d is outflow; face (i,j,k) of direction d is
// the face below cell (i,j,k).
struct Grid {
    int n[3];
    int per[3];
    int fe(int d, int q) const { return n[q] + (q == d && !per[q] ? 1 : 0); }
    long long flen(int d) const { return (long long)fe(d, 0) * fe(d, 1) * fe(d, 2); }
    long long cells() const { return (long long)n[0] * n[1] * n[2]; }
    int wrap(int x, int q) const { return ((x % n[q]) + n[q]) % n[q]; }
    // state read: periodic image, or the nearest valid cell (the foextrap
    // ghost fill, MaestroBCFill.cpp:8-171)
    long long at(int i, int j, int k) const {
        auto w = [&](int x, int q) {
            return per[q] ? wrap(x, q) : (x < 0 ? 0 : (x >= n[q] ? n[q] - 1 : x));
        };
        return (long long)w(i, 0) + (long long)n[0] * ((long long)w(j, 1) + (long long)n[1] * w(k, 2));
    }
    // face (i,j,k) of direction d; outflow positions must lie in the array
    long long face(int d, int i, int j, int k) const {
        const int p[3] = {i, j, k};
        long long r = 0, m = 1;
        for (int q = 0; q < 3; ++q) {
            int x = p[q];
            if (per[q]) {
                x = wrap(x, q);
            } else if (x < 0 || x >= fe(d, q)) {
                std::fprintf(stderr, "oracle step: face (%d,%d,%d) of direction %d out of range\n",
                             i, j, k, d);
                std::abort();
            }
            r += m * x;
            m *= fe(d, q);
        }
        return r;
    }
    // 1: d-face at the low outflow side, 2: at the high one, else 0
    int bnd(int d, int pos) const {
        if (per[d]) return 0;
        return pos == 0 ? 1 : (pos == n[d] ? 2 : 0);
    }
};

inline void step_e(int d, int& a, int& b, int& c) {
    a = d == 0;
    b = d == 1;
    c = d == 2;
}

// the boundary rule of a d-face at an outflow side, for the predictor,
// transverse and edge states alike (MaestroMakeEdgeScalars.cpp:709-738,
// :931-961, :1400-1430): the state from outside the domain takes the
// inside one (foextrap)
template <class T>
inline void outflow_face(int b, T& l, T& r) {
    if (b == 1) l = r;
    else if (b == 2) r = l;
}

// one step of one component; s and snew are cell arrays, u[d] face arrays
// marks (op counting only): the counter after each of the five phases
// cons: the conservative form of MakeEdgeScal (MaestroMakeEdgeScalars.cpp:
// 906-929 transverse, :1411-1438 edges, MakeDivU :648-660)
template <class T>
void step_comp(const Grid& g, const T* s, const T* const u[3], T* snew, double dt,
               const double dx[3], double rel_eps, long long* marks = nullptr, int ppm = 1,
               bool cons = false) {
    const long long NC = g.cells();
    auto cell = [&](int i, int j, int k) { return g.at(i, j, k); };
    std::vector<T> Ip[3], Im[3], simh[3], corner[6], se[3];
    auto pair_index = [](int d, int t) { return d * 2 + (t > d ? t - 1 : t); };
    for (int d = 0; d < 3; ++d) {
        Ip[d].resize(NC);
        Im[d].resize(NC);
        simh[d].resize(g.flen(d));
        se[d].resize(g.flen(d));
        for (int t = 0; t < 3; ++t)
            if (t != d) corner[pair_index(d, t)].resize(g.flen(d));
    }

    // PPM in every direction
    for (int k = 0; k < g.n[2]; ++k)
        for (int j = 0; j < g.n[1]; ++j)
            for (int i = 0; i < g.n[0]; ++i)
                for (int d = 0; d < 3; ++d) {
                    int a, b, c;
                    step_e(d, a, b, c);
                    T v[7];
                    for (int o = -3; o <= 3; ++o) v[o + 3] = s[cell(i + o * a, j + o * b, k + o * c)];
                    T sm, sp;
                    if (ppm == 1) {
                        ppm1(v + 3, sm, sp);
                    } else {
                        // faces cell-1 .. cell+2
                        const T fl = cs_face(v[0], v[1], v[2], v[3]);
                        const T f0 = cs_face(v[1], v[2], v[3], v[4]);
                        const T fr = cs_face(v[2], v[3], v[4], v[5]);
                        const T frr = cs_face(v[3], v[4], v[5], v[6]);
                        cs_cell(v + 3, fl, f0, fr, frr, sm, sp);
                    }
                    trace(v[3], sm, sp, u[d][g.face(d, i + a, j + b, k + c)], u[d][g.face(d, i, j, k)],
                          dt, dx[d], rel_eps, Ip[d][cell(i, j, k)], Im[d][cell(i, j, k)]);
                }
    if (marks) marks[0] = g_ops;
    // loops over the d-faces: f(d, i, j, k, b) with b the boundary side
    auto faces = [&](int d, auto&& f) {
        for (int k = 0; k < g.fe(d, 2); ++k)
            for (int j = 0; j < g.fe(d, 1); ++j)
                for (int i = 0; i < g.fe(d, 0); ++i) f(i, j, k, g.bnd(d, d == 0 ? i : (d == 1 ? j : k)));
    };
    // predictor and upwind states on faces (3-D threshold 0)
    for (int d = 0; d < 3; ++d) {
        int a, b, c;
        step_e(d, a, b, c);
        faces(d, [&](int i, int j, int k, int bd) {
            T l = cst<T>(0.0), r = cst<T>(0.0);
            if (bd != 1) l = Ip[d][cell(i - a, j - b, k - c)];
            if (bd != 2) r = Im[d][cell(i, j, k)];
            outflow_face(bd, l, r);
            const long long f = g.face(d, i, j, k);
            simh[d][f] = upwind(l, r, u[d][f], 0.0);
        });
    }
    if (marks) marks[1] = g_ops;
    const double dt2 = 0.5 * dt;
    // corner-coupled states simh_dt on d-faces, corrected by t: the
    // state `base` of cell (i,j,k) with its transverse term in direction t
    const double dt6 = dt / 6.0;
    const double dt3 = dt / 3.0;
    // MakeDivU of cell (i,j,k): the six face differences over dx[0]
    auto divu = [&](int i, int j, int k) {
        T dv = u[0][g.face(0, i + 1, j, k)] - u[0][g.face(0, i, j, k)] + u[1][g.face(1, i, j + 1, k)] -
               u[1][g.face(1, i, j, k)] + u[2][g.face(2, i, j, k + 1)] - u[2][g.face(2, i, j, k)];
        return dv / dx[0];
    };
    auto corrected = [&](T base, int t, int i, int j, int k) {
        int p, q, r;
        step_e(t, p, q, r);
        const long long f0 = g.face(t, i, j, k), f1 = g.face(t, i + p, j + q, k + r);
        if (!cons)
            return base - (dt6 / dx[t]) * (u[t][f1] + u[t][f0]) * (simh[t][f1] - simh[t][f0]);
        const T sc = s[cell(i, j, k)];
        return base - (dt3 / dx[t]) * (simh[t][f1] * u[t][f1] - simh[t][f0] * u[t][f0]) -
               dt3 * sc * divu(i, j, k) + (dt3 / dx[t]) * sc * (u[t][f1] - u[t][f0]);
    };
    for (int d = 0; d < 3; ++d)
        for (int t = 0; t < 3; ++t) {
            if (t == d) continue;
            int a, b, c;
            step_e(d, a, b, c);
            faces(d, [&](int i, int j, int k, int bd) {
                T l = cst<T>(0.0), rr = cst<T>(0.0);
                if (bd != 1) l = corrected(Ip[d][cell(i - a, j - b, k - c)], t, i - a, j - b, k - c);
                if (bd != 2) rr = corrected(Im[d][cell(i, j, k)], t, i, j, k);
                outflow_face(bd, l, rr);
                const long long f = g.face(d, i, j, k);
                corner[pair_index(d, t)][f] = upwind(l, rr, u[d][f], rel_eps);
            });
        }
    if (marks) marks[2] = g_ops;
    // edge states: d-face terms in the order of the reference (x: y, z;
    // y: x, z; z: x, y), force zero; the edge term of cell (i,j,k) from the
    // corner state on its t-faces corrected by o
    const double dt4 = 0.25 * dt;
    auto eterm = [&](int t, int o, int i, int j, int k) {
        int p, q, r;
        step_e(t, p, q, r);
        const std::vector<T>& ct = corner[pair_index(t, o)];
        const long long f0 = g.face(t, i, j, k), f1 = g.face(t, i + p, j + q, k + r);
        if (cons) return (dt2 / dx[t]) * (ct[f1] * u[t][f1] - ct[f0] * u[t][f0]);
        return (dt4 / dx[t]) * (u[t][f1] + u[t][f0]) * (ct[f1] - ct[f0]);
    };
    // the conservative form's normal term of cell (i,j,k) on its d-faces
    auto nterm = [&](int d, int i, int j, int k) {
        int a, b, c;
        step_e(d, a, b, c);
        return (dt2 / dx[d]) * s[cell(i, j, k)] *
               (u[d][g.face(d, i + a, j + b, k + c)] - u[d][g.face(d, i, j, k)]);
    };
    for (int d = 0; d < 3; ++d) {
        int a, b, c;
        step_e(d, a, b, c);
        faces(d, [&](int i, int j, int k, int bd) {
            T l = cst<T>(0.0), rr = cst<T>(0.0);
            if (bd != 1) {
                l = Ip[d][cell(i - a, j - b, k - c)];
                for (int t = 0; t < 3; ++t)
                    if (t != d) l = l - eterm(t, 3 - d - t, i - a, j - b, k - c);
                if (cons) l = l - nterm(d, i - a, j - b, k - c);
            }
            if (bd != 2) {
                rr = Im[d][cell(i, j, k)];
                for (int t = 0; t < 3; ++t)
                    if (t != d) rr = rr - eterm(t, 3 - d - t, i, j, k);
                if (cons) rr = rr - nterm(d, i, j, k);
            }
            T sedgel = l + dt2 * 0.0;
            T sedger = rr + dt2 * 0.0;
            outflow_face(bd, sedgel, sedger);
            const long long f = g.face(d, i, j, k);
            se[d][f] = upwind(sedgel, sedger, u[d][f], rel_eps);
        });
    }
    if (marks) marks[3] = g_ops;
    // fluxes and the update
    for (int k = 0; k < g.n[2]; ++k)
        for (int j = 0; j < g.n[1]; ++j)
            for (int i = 0; i < g.n[0]; ++i) {
                const long long cc = cell(i, j, k);
                auto fl = [&](int d, int i2, int j2, int k2) {
                    const long long f = g.face(d, i2, j2, k2);
                    return u[d][f] * se[d][f];
                };
                T fx1 = fl(0, i + 1, j, k);
                T fx0 = fl(0, i, j, k);
                T fy1 = fl(1, i, j + 1, k);
                T fy0 = fl(1, i, j, k);
                T fz1 = fl(2, i, j, k + 1);
                T fz0 = fl(2, i, j, k);
                T divterm = (fx1 - fx0) / dx[0];
                divterm = divterm + (fy1 - fy0) / dx[1];
                divterm = divterm + (fz1 - fz0) / dx[2];
                snew[cc] = s[cc] + dt * (-divterm + 0.0);
            }
    if (marks) marks[4] = g_ops;
}

}  // namespace

// FP64 operations per cell of one component of the step, in the per-point
// formulation of the reference (every cell evaluates its own three van Leer
// slopes and two face values per direction), on the data given:
// out[0..4] the phases PPM and trace, predictor upwinding, corner-coupled
// states, edge states, flux and update; out[5] one van Leer slope and out[6]
// one PPM type 1 face value (limits included), averaged over the calls.
void StepOpCount(int n, double dt, const double dx[3], double rel_eps, const double* s,
                 const double* const u[3], double out[7]) {
    const Grid g{{n, n, n}, {1, 1, 1}};
    const long long N3 = (long long)n * n * n;
    std::vector<Cnt> sc(N3), out_c(N3), uc[3];
    const Cnt* up[3];
    for (long long x = 0; x < N3; ++x) sc[x] = {s[x]};
    for (int d = 0; d < 3; ++d) {
        uc[d].resize(N3);
        for (long long x = 0; x < N3; ++x) uc[d][x] = {u[d][x]};
        up[d] = uc[d].data();
    }
    long long marks[5];
    g_ops = 0;
    step_comp<Cnt>(g, sc.data(), up, out_c.data(), dt, dx, rel_eps, marks);
    long long prev = 0;
    for (int q = 0; q < 5; ++q) {
        out[q] = (double)(marks[q] - prev) / (double)N3;
        prev = marks[q];
    }
    // one slope and one face value, over every cell and direction
    long long slopes = 0, faces = 0;
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int d = 0; d < 3; ++d) {
                    int a, b, c;
                    step_e(d, a, b, c);
                    Cnt v[4];
                    for (int o = -2; o <= 1; ++o) v[o + 2] = sc[g.at(i + o * a, j + o * b, k + o * c)];
                    g_ops = 0;
                    const Cnt dl = vl(v[0], v[1], v[2]);
                    slopes += g_ops;
                    const Cnt dr = vl(v[1], v[2], v[3]);
                    g_ops = 0;
                    Cnt e = 0.5 * (v[2] + v[1]) - (dr - dl) / 6.0;
                    e = tmax(e, tmin(v[2], v[1]));
                    e = tmin(e, tmax(v[2], v[1]));
                    faces += g_ops;
                }
    out[5] = (double)slopes / (double)(3 * N3);
    out[6] = (double)faces / (double)(3 * N3);
}

// the step (T = double) or its directional derivative (ds, du != null) on
// n[0] x n[1] x n[2] cells, direction d periodic (per[d]) or outflow, in
// the advective or (cons) the conservative form:
// s, snew: ncomp cell arrays; u[d]: the face array of direction d (one more
// point along d when d is outflow; du[d] may be null)
void ScalStepBC(const int n[3], const int per[3], int ncomp, int ppm, double dt,
                const double dx[3], double rel_eps, const double* s, const double* const u[3],
                double* snew, const double* ds, const double* const du[3], double* dsnew,
                bool cons) {
    const Grid g{{n[0], n[1], n[2]}, {per[0] ? 1 : 0, per[1] ? 1 : 0, per[2] ? 1 : 0}};
    const long long NC = g.cells();
    // components are independent: one OpenMP thread each (the double and
    // dual instantiations keep no shared state)
    if (!ds) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int c = 0; c < ncomp; ++c)
            step_comp<double>(g, s + c * NC, u, snew + c * NC, dt, dx, rel_eps, nullptr, ppm,
                              cons);
        return;
    }
    std::vector<Dual> ud[3];
    const Dual* up[3];
    for (int d = 0; d < 3; ++d) {
        const long long nf = g.flen(d);
        ud[d].resize(nf);
        for (long long x = 0; x < nf; ++x) ud[d][x] = {u[d][x], du[d] ? du[d][x] : 0.0};
        up[d] = ud[d].data();
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int c = 0; c < ncomp; ++c) {
        std::vector<Dual> sd(NC), out(NC);
        for (long long x = 0; x < NC; ++x) sd[x] = {s[c * NC + x], ds[c * NC + x]};
        step_comp<Dual>(g, sd.data(), up, out.data(), dt, dx, rel_eps, nullptr, ppm, cons);
        for (long long x = 0; x < NC; ++x) {
            snew[c * NC + x] = out[x].v;
            dsnew[c * NC + x] = out[x].d;
        }
    }
}

// the step on a periodic N^3 box: s, snew: ncomp x N^3; u[d]: N^3 (face N
// is face 0)
void ScalStep(int n, int ncomp, int ppm, double dt, const double dx[3], double rel_eps,
              const double* s, const double* const u[3], double* snew, const double* ds,
              const double* const du[3], double* dsnew) {
    const int nn[3] = {n, n, n}, per[3] = {1, 1, 1};
    ScalStepBC(nn, per, ncomp, ppm, dt, dx, rel_eps, s, u, snew, ds, du, dsnew, false);
}

}  // namespace mxo
