// oracle_core.H -- containers for the CPU oracle of the advection path.
//
// TEST INFRASTRUCTURE ONLY.  The oracle is the checker the HIP library is
// compared against; nothing in the shipped library includes, links or calls
// it.  It restates the reference's CPU algorithm (AMReX ParallelFor bodies
// run over boxes with OpenMP) on plain C++ containers: an index box with
// grow/growLo/growHi, an Array4-like view (i fastest, component outermost),
// and MultiFab temporaries allocated per call as the reference does.
#ifndef MXO_ORACLE_CORE_H
#define MXO_ORACLE_CORE_H

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <vector>

#include "mx_adv.h"

namespace mxo {

// amrex::min / amrex::max are std::min / std::max: the first argument wins
// ties (relevant for signed zeros); amrex::Math::abs / copysign are fabs /
// copysign.
inline double mn(double a, double b) { return (b < a) ? b : a; }
inline double mx(double a, double b) { return (a < b) ? b : a; }
inline double ab(double a) { return std::fabs(a); }
inline double cs(double a, double b) { return std::copysign(a, b); }

extern std::atomic<int> g_oob;  // set when a bounds-checked access misses
// wall time spent allocating and initialising temporaries (Fab::define), so
// that a caller can report allocation-inclusive and -exclusive times
extern std::atomic<long long> g_alloc_ns;

struct IBox {
    int lo[3];
    int hi[3];
};

inline IBox make_box(const int* lo, const int* hi) {
    IBox b;
    for (int d = 0; d < 3; ++d) {
        b.lo[d] = lo[d];
        b.hi[d] = hi[d];
    }
    return b;
}

// amrex::grow(bx, n) in the first `dim` directions
inline IBox grow(IBox b, int dim, int n) {
    for (int d = 0; d < dim; ++d) {
        b.lo[d] -= n;
        b.hi[d] += n;
    }
    return b;
}
inline IBox grow_dir(IBox b, int d, int n) {
    b.lo[d] -= n;
    b.hi[d] += n;
    return b;
}
inline IBox grow_lo(IBox b, int d, int n) {
    b.lo[d] -= n;
    return b;
}
inline IBox grow_hi(IBox b, int d, int n) {
    b.hi[d] += n;
    return b;
}
// surroundingNodes(bx, d): the faces normal to d
inline IBox nodal(IBox b, int d) {
    b.hi[d] += 1;
    return b;
}

struct Arr {
    double* p = nullptr;
    long long js = 0, ks = 0, ns = 0;
    int b[3] = {0, 0, 0};
    int e[3] = {0, 0, 0};
    int nc = 0;

    double& operator()(int i, int j, int k, int n = 0) const {
#ifdef MXO_BOUNDS
        if (i < b[0] || i >= e[0] || j < b[1] || j >= e[1] || k < b[2] ||
            k >= e[2] || n < 0 || n >= nc || p == nullptr) {
            g_oob.store(1);
            static thread_local double dummy;
            dummy = 0.0;
            return dummy;
        }
#endif
        return p[(i - b[0]) + (j - b[1]) * js + (k - b[2]) * ks + n * ns];
    }
    bool valid() const { return p != nullptr; }
};

inline Arr from_fab(const mx_fab& f) {
    Arr a;
    a.p = f.p;
    a.js = f.jstride;
    a.ks = f.kstride;
    a.ns = f.nstride;
    for (int d = 0; d < 3; ++d) {
        a.b[d] = f.begin[d];
        a.e[d] = f.end[d];
    }
    a.nc = f.ncomp;
    return a;
}

// owned FAB storage (a temporary MultiFab's data for one box)
struct Fab {
    std::vector<double> v;
    Arr a;
    void define(const IBox& region, int ncomp, double init = 0.0) {
        const auto t0 = std::chrono::steady_clock::now();
        long long n0 = region.hi[0] - region.lo[0] + 1;
        long long n1 = region.hi[1] - region.lo[1] + 1;
        long long n2 = region.hi[2] - region.lo[2] + 1;
        v.assign((size_t)(n0 * n1 * n2 * ncomp), init);
        a.p = v.data();
        a.js = n0;
        a.ks = n0 * n1;
        a.ns = n0 * n1 * n2;
        for (int d = 0; d < 3; ++d) {
            a.b[d] = region.lo[d];
            a.e[d] = region.hi[d] + 1;
        }
        a.nc = ncomp;
        g_alloc_ns += std::chrono::duration_cast<std::chrono::nanoseconds>(
                          std::chrono::steady_clock::now() - t0)
                          .count();
    }
};

// One level of a temporary MultiFab: FAB per box, `ng` ghosts, cell- or
// face-centred (nodal_dir >= 0).
struct TmpMF {
    std::vector<Fab> fab;
    void define(const std::vector<IBox>& grids, int dim, int nodal_dir,
                int ncomp, int ng, double init = 0.0) {
        fab.resize(grids.size());
        for (size_t b = 0; b < grids.size(); ++b) {
            IBox r = grids[b];
            if (nodal_dir >= 0) r = nodal(r, nodal_dir);
            r = grow(r, dim, ng);
            fab[b].define(r, ncomp, init);
        }
    }
    const Arr& operator[](size_t b) const { return fab[b].a; }
};

inline std::vector<Arr> arrs(const mx_multifab* m) {
    std::vector<Arr> v;
    if (!m || !m->fab) return v;
    for (int b = 0; b < m->nfab; ++b) v.push_back(from_fab(m->fab[b]));
    return v;
}

// ParallelFor over a box: OpenMP over the (k, j) planes, i innermost.
template <class F>
inline void par_for(const IBox& bx, F&& f) {
    const int klo = bx.lo[2], khi = bx.hi[2];
    const int jlo = bx.lo[1], jhi = bx.hi[1];
    const int ilo = bx.lo[0], ihi = bx.hi[0];
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = klo; k <= khi; ++k)
        for (int j = jlo; j <= jhi; ++j)
            for (int i = ilo; i <= ihi; ++i) f(i, j, k);
}

struct BaseArr {
    const double* p = nullptr;
    int len = 0;
    double operator()(int lev, int r) const { return p[(long long)lev * len + r]; }
};
inline BaseArr base(const mx_basestate* s) {
    BaseArr a;
    if (s) {
        a.p = s->p;
        a.len = s->len;
    }
    return a;
}

// amrex::FluxRegister of level lev (the lev / lev - 1 interface): one face
// box per (fine box b, orientation o), o = d for the low side of direction d
// and dim + d for the high side (amrex::Orientation's order); the face box is
// in the coarse level's face index space, one face thick.
struct FluxReg {
    int ncomp = 0;
    std::vector<IBox> box;                 // [b * 2 * dim + o]
    std::vector<std::vector<double>> val;  // same index; i fastest, comp outermost
};

struct Ctx {
    mx_adv_params p;
    std::vector<IBox> grids[MX_MAX_LEV];
    // the flux registers are updated inside the flux stages (as the
    // reference's MakeRhoXFlux / MakeRhoHFlux do), which see the Ctx as const
    mutable FluxReg freg[MX_MAX_LEV];
    int dim() const { return p.dim; }
    IBox domain(int lev) const {
        IBox b = make_box(p.domain_lo[lev], p.domain_hi[lev]);
        if (p.dim == 2) b.lo[2] = b.hi[2] = 0;
        return b;
    }
};

}  // namespace mxo

#endif
