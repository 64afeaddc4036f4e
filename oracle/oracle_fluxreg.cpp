// oracle_fluxreg.cpp -- CPU restatement of the flux register Maestro uses
// with reflux_type == 2: FluxRegister::CrseInit and FineAdd inside
// MakeRhoXFlux / MakeRhoHFlux (MaestroMakeFlux.cpp:420-456, 1035-1060) and
// FluxRegister::Reflux inside UpdateScal (MaestroAdvection.cpp:258-278).
//
// TEST INFRASTRUCTURE ONLY (see oracle_core.H).  amrex::FluxRegister lives
// in AMReX (AMReX_FluxRegister, AMReX_FluxReg_C.H), an un-vendored submodule
// of the reference (external/amrex is empty and its commit is not recorded).
// This restates its published algorithm:
//   * define(fine grids, ratio, ncomp): per fine box and orientation, the
//     faces of the coarsened box on that side (bdryLo: face index lo[d];
//     bdryHi: hi[d] + 1), set to 0;
//   * CrseInit(flux_d, d, scomp, dcomp, n, mult): every register face of
//     direction d that a coarse box holds takes flux * mult (a copy; where two
//     coarse boxes hold the face, the first in box order gives it);
//   * FineAdd(flux_d, d, scomp, dcomp, n, mult): every register face of fine
//     box b adds mult * flux over the ratio^(dim-1) fine faces on it, one
//     addition per fine face, the fine face loops running (k, j) for d = 0,
//     (k, i) for d = 1 and (j, i) for d = 2, the first index outermost;
//   * Reflux(state, scale, scomp, dcomp, n): for each orientation in order
//     (the low sides of x, y, z, then the high sides), a face-centred flux on
//     the coarse boxes that is the register value where a register face lands
//     on the face (periodic images included) and 0 elsewhere, and for every
//     valid coarse cell s += (-scale * f(cell's high face)) / vol on a low
//     orientation, s += (scale * f(cell's low face)) / vol on a high one, vol
//     = dx[0] * dx[1] (* dx[2]) of the coarse level.
#include <algorithm>
#include <array>

#include "oracle_core.H"

namespace mxo {

namespace {

int reg_ratio(const Ctx& c, int lev, int d) {
    const IBox a = c.domain(lev - 1), b = c.domain(lev);
    return (b.hi[d] - b.lo[d] + 1) / (a.hi[d] - a.lo[d] + 1);
}

int fdiv(int a, int r) { return a >= 0 ? a / r : -((-a + r - 1) / r); }

long long npts(const IBox& b) {
    long long n = 1;
    for (int d = 0; d < 3; ++d) n *= (long long)(b.hi[d] - b.lo[d] + 1);
    return n;
}

bool inside(const IBox& b, int i, int j, int k) {
    return i >= b.lo[0] && i <= b.hi[0] && j >= b.lo[1] && j <= b.hi[1] && k >= b.lo[2] &&
           k <= b.hi[2];
}

long long pidx(const IBox& b, int i, int j, int k) {
    const long long nx = b.hi[0] - b.lo[0] + 1, ny = b.hi[1] - b.lo[1] + 1;
    return (i - b.lo[0]) + nx * ((j - b.lo[1]) + ny * (long long)(k - b.lo[2]));
}

}  // namespace

int FluxRegDefine(const Ctx& c, int lev, int ncomp) {
    if (lev < 1 || lev > c.p.finest_level || c.grids[lev].empty() || ncomp < 1) return MX_ERR_ARG;
    const int dim = c.dim();
    FluxReg& R = c.freg[lev];
    R.ncomp = ncomp;
    R.box.clear();
    R.val.clear();
    int r[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) r[d] = reg_ratio(c, lev, d);
    for (const IBox& fb : c.grids[lev]) {
        IBox cb;
        for (int d = 0; d < 3; ++d) {
            cb.lo[d] = fdiv(fb.lo[d], r[d]);
            cb.hi[d] = fdiv(fb.hi[d], r[d]);
        }
        for (int o = 0; o < 2 * dim; ++o) {
            const int d = o % dim;
            IBox f = cb;
            if (o < dim)
                f.hi[d] = f.lo[d];
            else
                f.lo[d] = f.hi[d] = cb.hi[d] + 1;
            R.box.push_back(f);
            R.val.emplace_back((size_t)(npts(f) * ncomp), 0.0);
        }
    }
    return MX_OK;
}

int FluxRegSetVal(const Ctx& c, int lev, double v) {
    if (lev < 1 || lev >= MX_MAX_LEV || c.freg[lev].ncomp == 0) return MX_ERR_ARG;
    for (auto& a : c.freg[lev].val) std::fill(a.begin(), a.end(), v);
    return MX_OK;
}

int FluxRegCrseInit(const Ctx& c, int lev, const std::vector<Arr>& flux, int dir, int scomp,
                    int dcomp, int ncomp, double mult) {
    FluxReg& R = c.freg[lev];
    const int dim = c.dim();
    const std::vector<IBox>& cg = c.grids[lev - 1];
    for (size_t q = 0; q < R.box.size(); ++q) {
        const int o = (int)(q % (2 * dim));
        if (o % dim != dir) continue;
        const IBox& f = R.box[q];
        std::vector<double>& v = R.val[q];
        const long long np = npts(f);
        for (int k = f.lo[2]; k <= f.hi[2]; ++k)
            for (int j = f.lo[1]; j <= f.hi[1]; ++j)
                for (int i = f.lo[0]; i <= f.hi[0]; ++i) {
                    for (size_t b = 0; b < cg.size(); ++b) {
                        if (!inside(nodal(cg[b], dir), i, j, k)) continue;
                        for (int n = 0; n < ncomp; ++n)
                            v[(dcomp + n) * np + pidx(f, i, j, k)] = flux[b](i, j, k, scomp + n) * mult;
                        break;
                    }
                }
    }
    return MX_OK;
}

int FluxRegFineAdd(const Ctx& c, int lev, const std::vector<Arr>& flux, int dir, int scomp,
                   int dcomp, int ncomp, double mult) {
    FluxReg& R = c.freg[lev];
    const int dim = c.dim();
    int r[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) r[d] = reg_ratio(c, lev, d);
    for (size_t q = 0; q < R.box.size(); ++q) {
        const int o = (int)(q % (2 * dim));
        if (o % dim != dir) continue;
        const Arr& fa = flux[q / (2 * dim)];
        const IBox& f = R.box[q];
        std::vector<double>& v = R.val[q];
        const long long np = npts(f);
        for (int n = 0; n < ncomp; ++n)
            for (int k = f.lo[2]; k <= f.hi[2]; ++k)
                for (int j = f.lo[1]; j <= f.hi[1]; ++j)
                    for (int i = f.lo[0]; i <= f.hi[0]; ++i) {
                        double& x = v[(dcomp + n) * np + pidx(f, i, j, k)];
                        const int sc = scomp + n;
                        if (dir == 0) {
                            for (int ko = 0; ko < r[2]; ++ko)
                                for (int jo = 0; jo < r[1]; ++jo)
                                    x += mult * fa(i * r[0], j * r[1] + jo, k * r[2] + ko, sc);
                        } else if (dir == 1) {
                            for (int ko = 0; ko < r[2]; ++ko)
                                for (int io = 0; io < r[0]; ++io)
                                    x += mult * fa(i * r[0] + io, j * r[1], k * r[2] + ko, sc);
                        } else {
                            for (int jo = 0; jo < r[1]; ++jo)
                                for (int io = 0; io < r[0]; ++io)
                                    x += mult * fa(i * r[0] + io, j * r[1] + jo, k * r[2], sc);
                        }
                    }
    }
    return MX_OK;
}

int FluxRegReflux(const Ctx& c, int lev, const std::vector<Arr>& state, double scale, int scomp,
                  int dcomp, int ncomp) {
    const FluxReg& R = c.freg[lev];
    const int dim = c.dim();
    const IBox dom = c.domain(lev - 1);
    const double* dx = c.p.dx[lev - 1];
    const double vol = dim == 3 ? dx[0] * dx[1] * dx[2] : dx[0] * dx[1];
    // periodic shifts of the coarse domain (0 included)
    std::vector<std::array<int, 3>> shifts;
    for (int a = -1; a <= 1; ++a)
        for (int b = -1; b <= 1; ++b)
            for (int e = -1; e <= 1; ++e) {
                const int s[3] = {a, b, e};
                bool ok = true;
                for (int d = 0; d < 3; ++d)
                    if (s[d] != 0 && (d >= dim || !c.p.is_periodic[d])) ok = false;
                if (!ok) continue;
                std::array<int, 3> sh;
                for (int d = 0; d < 3; ++d) sh[d] = s[d] * (dom.hi[d] - dom.lo[d] + 1);
                shifts.push_back(sh);
            }
    const std::vector<IBox>& cg = c.grids[lev - 1];
    for (size_t b = 0; b < cg.size(); ++b) {
        const Arr& s = state[b];
        par_for(cg[b], [&](int i, int j, int k) {
            for (int n = 0; n < ncomp; ++n) {
                double x = s(i, j, k, dcomp + n);
                for (int o = 0; o < 2 * dim; ++o) {
                    const int d = o % dim;
                    const bool lo = o < dim;
                    int F[3] = {i, j, k};
                    if (lo) F[d] += 1;  // the cell's high face
                    double f = 0.0;
                    bool found = false;
                    for (size_t q = o; q < R.box.size() && !found; q += 2 * dim)
                        for (const auto& sh : shifts) {
                            const int fi = F[0] - sh[0], fj = F[1] - sh[1], fk = F[2] - sh[2];
                            if (!inside(R.box[q], fi, fj, fk)) continue;
                            f = R.val[q][(scomp + n) * npts(R.box[q]) + pidx(R.box[q], fi, fj, fk)];
                            found = true;
                            break;
                        }
                    if (lo)
                        x += -scale * f / vol;
                    else
                        x += scale * f / vol;
                }
                s(i, j, k, dcomp + n) = x;
            }
        });
    }
    return MX_OK;
}

// the register updates of MakeRhoXFlux (rhox: components start_comp ..
// start_comp + num_comp - 1 and Rho) or MakeRhoHFlux (RhoH) on level lev with
// reflux_type 2 (MaestroMakeFlux.cpp:420-451, 1035-1058): CrseInit into the
// register of lev + 1 with -1.0 * dt * area[i], FineAdd into the register of
// lev with 1.0 * dt * area[i], area the level's face areas, direction by
// direction
int FluxRegOps(const Ctx& c, int lev, const std::vector<Arr> sflux[3], bool rhox,
               int start_comp, int num_comp) {
    const mx_adv_params& p = c.p;
    if (p.reflux_type != 2) return MX_OK;
    const int dim = p.dim;
    const double* dx = p.dx[lev];
    double area[3] = {0.0, 0.0, 0.0};
    if (dim == 3) {
        area[0] = dx[1] * dx[2];
        area[1] = dx[0] * dx[2];
        area[2] = dx[0] * dx[1];
    } else {
        area[0] = dx[1];
        area[1] = dx[0];
    }
    const double dt = p.dt;
    const bool crse = lev + 1 <= p.finest_level && lev + 1 < MX_MAX_LEV && c.freg[lev + 1].ncomp > 0;
    const bool fine = lev >= 1 && c.freg[lev].ncomp > 0;
    const int need = rhox ? std::max(start_comp + num_comp, p.comp_rho + 1) : p.comp_rhoh + 1;
    if ((crse && c.freg[lev + 1].ncomp < need) || (fine && c.freg[lev].ncomp < need))
        return MX_ERR_ARG;
    for (int i = 0; i < dim && crse; ++i) {
        const double m = -1.0 * dt * area[i];
        if (rhox) {
            FluxRegCrseInit(c, lev + 1, sflux[i], i, start_comp, start_comp, num_comp, m);
            FluxRegCrseInit(c, lev + 1, sflux[i], i, p.comp_rho, p.comp_rho, 1, m);
        } else {
            FluxRegCrseInit(c, lev + 1, sflux[i], i, p.comp_rhoh, p.comp_rhoh, 1, m);
        }
    }
    for (int i = 0; i < dim && fine; ++i) {
        const double m = 1.0 * dt * area[i];
        if (rhox) {
            FluxRegFineAdd(c, lev, sflux[i], i, start_comp, start_comp, num_comp, m);
            FluxRegFineAdd(c, lev, sflux[i], i, p.comp_rho, p.comp_rho, 1, m);
        } else {
            FluxRegFineAdd(c, lev, sflux[i], i, p.comp_rhoh, p.comp_rhoh, 1, m);
        }
    }
    return MX_OK;
}

}  // namespace mxo
