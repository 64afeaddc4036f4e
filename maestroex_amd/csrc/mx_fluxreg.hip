// mx_fluxreg.hip -- the flux register of reflux_type 2 (amrex::FluxRegister
// as Maestro uses it): CrseInit and FineAdd inside MakeRhoXFlux /
// MakeRhoHFlux (MaestroMakeFlux.cpp:420-456, 1035-1060) and Reflux inside
// UpdateScal (MaestroAdvection.cpp:258-278).
//
// The register of level lev (the lev / lev - 1 interface) holds, per fine box
// b and orientation o (o = d: the low side of direction d, o = dim + d: the
// high side, amrex::Orientation's order), the faces of the coarsened box on
// that side in the coarse level's face index space (face lo[d], or hi[d] + 1),
// one FP64 array per register, components outermost, i fastest.  All
// registers of a level are one device allocation.
//
//   CrseInit(flux_d, d, scomp, dcomp, n, mult): a register face of direction
//     d takes flux * mult from the first coarse box (in box order) whose
//     valid faces hold it.
//   FineAdd(flux_d, d, scomp, dcomp, n, mult): a register face of fine box b
//     adds mult * flux of each of the ratio^(dim-1) fine faces on it, one
//     addition per fine face, the fine face loops running (k, j) for d = 0,
//     (k, i) for d = 1, (j, i) for d = 2, the first outermost.
//   Reflux(state, scale, scomp, dcomp, n): per valid coarse cell, for each
//     orientation in order, f = the register value landing on the cell's face
//     on the far side of that orientation (its high face for a low orientation,
//     its low face for a high one; periodic images included; 0 where none),
//     and s += (-scale * f) / vol (low) or s += (scale * f) / vol (high), vol =
//     the coarse cell volume dx[0] * dx[1] (* dx[2]).
#include <algorithm>
#include <array>
#include <cstring>
#include <vector>

#include "mx_internal.H"

using namespace mxd;
using namespace mxh;

namespace {

constexpr int NTF = 256;

struct FrInit {     // a register of the launch's direction (CrseInit)
    long long off;  // value offset of the register
    int lo[3], hi[3];
    int cand_off, cand_n;
};
struct FrCand {  // a coarse box's valid faces of the direction
    DA src;
    int lo[3], hi[3];
};
struct FrAdd {  // a register of the launch's direction (FineAdd)
    long long off;
    int lo[3], hi[3];
    DA fine;  // the fine box's flux of that direction
};
struct FrCell {  // a coarse box (Reflux)
    DA s;
    int lo[3], hi[3];
    int ref_off[7];  // refs of orientation o: [ref_off[o], ref_off[o + 1])
};
struct FrRef {  // a register (or periodic image) near the coarse box
    long long off;
    int lo[3], hi[3];  // the register's faces, shifted
    int pad_;
};

__device__ __forceinline__ long long fr_idx(const int* lo, const int* hi, int i, int j, int k) {
    const long long nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1;
    return (i - lo[0]) + nx * ((j - lo[1]) + ny * (long long)(k - lo[2]));
}
__device__ __forceinline__ long long fr_np(const int* lo, const int* hi) {
    return (long long)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
}
__device__ __forceinline__ bool fr_in(const int* lo, const int* hi, int i, int j, int k) {
    return i >= lo[0] && i <= hi[0] && j >= lo[1] && j <= hi[1] && k >= lo[2] && k <= hi[2];
}

__global__ void k_fr_fill(double* __restrict__ p, long long n, double v) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

__global__ void __launch_bounds__(NTF)
    k_fr_crse_init(const FrInit* __restrict__ regs, const FrCand* __restrict__ cands,
                   double* __restrict__ val, int scomp, int dcomp, int ncomp, double mult) {
    const FrInit& R = regs[blockIdx.y];
    int i, j, k;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (!region_point(t, R.lo, R.hi, i, j, k)) return;
    const long long np = fr_np(R.lo, R.hi);
    for (int q = 0; q < R.cand_n; ++q) {
        const FrCand& C = cands[R.cand_off + q];
        if (!fr_in(C.lo, C.hi, i, j, k)) continue;
        for (int n = 0; n < ncomp; ++n) val[R.off + (dcomp + n) * np + t] = C.src(i, j, k, scomp + n) * mult;
        return;
    }
}

__global__ void __launch_bounds__(NTF)
    k_fr_fine_add(const FrAdd* __restrict__ regs, double* __restrict__ val, int dir, int r0, int r1,
                  int r2, int scomp, int dcomp, int ncomp, double mult) {
    const FrAdd& R = regs[blockIdx.y];
    int i, j, k;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (!region_point(t, R.lo, R.hi, i, j, k)) return;
    const long long np = fr_np(R.lo, R.hi);
    for (int n = 0; n < ncomp; ++n) {
        double x = val[R.off + (dcomp + n) * np + t];
        const int sc = scomp + n;
        if (dir == 0) {
            for (int ko = 0; ko < r2; ++ko)
                for (int jo = 0; jo < r1; ++jo) x += mult * R.fine(i * r0, j * r1 + jo, k * r2 + ko, sc);
        } else if (dir == 1) {
            for (int ko = 0; ko < r2; ++ko)
                for (int io = 0; io < r0; ++io) x += mult * R.fine(i * r0 + io, j * r1, k * r2 + ko, sc);
        } else {
            for (int jo = 0; jo < r1; ++jo)
                for (int io = 0; io < r0; ++io) x += mult * R.fine(i * r0 + io, j * r1 + jo, k * r2, sc);
        }
        val[R.off + (dcomp + n) * np + t] = x;
    }
}

__global__ void __launch_bounds__(NTF)
    k_fr_reflux(const FrCell* __restrict__ cells, const FrRef* __restrict__ refs,
                const double* __restrict__ val, int dim, double scale, double vol, int scomp,
                int dcomp, int ncomp) {
    const FrCell& B = cells[blockIdx.y];
    int i, j, k;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (!region_point(t, B.lo, B.hi, i, j, k)) return;
    for (int n = 0; n < ncomp; ++n) {
        double x = B.s(i, j, k, dcomp + n);
        for (int o = 0; o < 2 * dim; ++o) {
            const int d = o < dim ? o : o - dim;
            const bool lo = o < dim;
            const int fi = i + (lo && d == 0), fj = j + (lo && d == 1), fk = k + (lo && d == 2);
            double f = 0.0;
            for (int q = B.ref_off[o]; q < B.ref_off[o + 1]; ++q) {
                const FrRef& R = refs[q];
                if (!fr_in(R.lo, R.hi, fi, fj, fk)) continue;
                f = val[R.off + (scomp + n) * fr_np(R.lo, R.hi) + fr_idx(R.lo, R.hi, fi, fj, fk)];
                break;
            }
            if (lo)
                x += -scale * f / vol;
            else
                x += scale * f / vol;
        }
        B.s(i, j, k, dcomp + n) = x;
    }
}

int fr_ratio(const mx_adv_ctx* c, int lev, int r[3]) {
    for (int d = 0; d < 3; ++d) r[d] = 1;
    for (int d = 0; d < c->p.dim; ++d) {
        const int nc = c->p.domain_hi[lev - 1][d] - c->p.domain_lo[lev - 1][d] + 1;
        const int nf = c->p.domain_hi[lev][d] - c->p.domain_lo[lev][d] + 1;
        if (nc <= 0 || nf % nc != 0) return 0;
        r[d] = nf / nc;
    }
    return 1;
}

int fdiv(int a, int r) { return a >= 0 ? a / r : -((-a + r - 1) / r); }

bool hit(const HBox& a, const HBox& b) {
    for (int d = 0; d < 3; ++d)
        if (std::max(a.lo[d], b.lo[d]) > std::min(a.hi[d], b.hi[d])) return false;
    return true;
}

void set_lohi(int* lo, int* hi, const HBox& b) {
    for (int d = 0; d < 3; ++d) {
        lo[d] = b.lo[d];
        hi[d] = b.hi[d];
    }
}

int fr_check(mx_adv_ctx* c, int lev) {
    if (!c || lev < 1 || lev > c->p.finest_level || lev >= MX_MAX_LEV) return MX_ERR_ARG;
    if (c->freg[lev].ncomp == 0 || !c->freg[lev].d) return MX_ERR_ARG;
    return MX_OK;
}

// one direction's flux MultiFab on level lev: a FAB per box holding its
// valid faces and components [0, nc)
int check_dir_flux(const mx_adv_ctx* c, int lev, const mx_multifab* m, int dir, int nc) {
    if (!m || !m->fab || m->nfab != (int)c->grids[lev].size()) return MX_ERR_ARG;
    for (int b = 0; b < m->nfab; ++b)
        if (!covers(m->fab[b], nodal(hbox(c->grids[lev][b]), dir), nc)) return MX_ERR_BOX;
    return MX_OK;
}

}  // namespace

namespace mxh {

void flux_reg_free(mx_adv_ctx* c, int lev) {
    mx_adv_ctx::FluxRegDev& R = c->freg[lev];
    if (R.d) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(R.d);
    }
    R = mx_adv_ctx::FluxRegDev();
}

int flux_reg_define(mx_adv_ctx* c, int lev, int ncomp) {
    if (!c || lev < 1 || lev > c->p.finest_level || lev >= MX_MAX_LEV || ncomp < 1 ||
        c->grids[lev].empty())
        return MX_ERR_ARG;
    int r[3];
    if (!fr_ratio(c, lev, r)) return MX_ERR_ARG;
    flux_reg_free(c, lev);
    mx_adv_ctx::FluxRegDev& R = c->freg[lev];
    const int dim = c->p.dim;
    long long total = 0;
    for (const mx_box& g : c->grids[lev]) {
        HBox cb = hbox(g);
        for (int d = 0; d < 3; ++d) {
            cb.lo[d] = fdiv(g.lo[d], r[d]);
            cb.hi[d] = fdiv(g.hi[d], r[d]);
        }
        for (int o = 0; o < 2 * dim; ++o) {
            const int d = o % dim;
            HBox f = cb;
            if (o < dim)
                f.hi[d] = f.lo[d];
            else
                f.lo[d] = f.hi[d] = cb.hi[d] + 1;
            mx_box fb;
            set_lohi(fb.lo, fb.hi, f);
            R.box.push_back(fb);
            R.off.push_back(total);
            total += npts(f) * ncomp;
        }
    }
    if (hipMalloc((void**)&R.d, (size_t)total * sizeof(double)) != hipSuccess) {
        R = mx_adv_ctx::FluxRegDev();
        return MX_ERR_NOMEM;
    }
    R.ncomp = ncomp;
    R.total = total;
    MX_LAUNCH(c, "k_fr_fill", k_fr_fill, dim3((unsigned)((total + NTF - 1) / NTF)), dim3(NTF), R.d,
              total, 0.0);
    return MX_OK;
}

int flux_reg_crse_init(mx_adv_ctx* c, int lev, const mx_multifab* flux, int dir, int scomp,
                       int dcomp, int ncomp, double mult) {
    mx_adv_ctx::FluxRegDev& R = c->freg[lev];
    const int dim = c->p.dim;
    const auto& cg = c->grids[lev - 1];
    std::vector<FrCand> cands;
    std::vector<HBox> cfaces;
    for (size_t b = 0; b < cg.size(); ++b) {
        FrCand C;
        C.src = to_da(flux->fab[b]);
        cfaces.push_back(nodal(hbox(cg[b]), dir));
        set_lohi(C.lo, C.hi, cfaces.back());
        cands.push_back(C);
    }
    std::vector<FrInit> regs;
    std::vector<FrCand> list;
    long long maxp = 0;
    for (size_t q = 0; q < R.box.size(); ++q) {
        if ((int)(q % (2 * dim)) % dim != dir) continue;
        const HBox f = hbox(R.box[q]);
        FrInit I;
        I.off = R.off[q];
        set_lohi(I.lo, I.hi, f);
        I.cand_off = (int)list.size();
        for (size_t b = 0; b < cands.size(); ++b)
            if (hit(f, cfaces[b])) list.push_back(cands[b]);
        I.cand_n = (int)list.size() - I.cand_off;
        maxp = std::max(maxp, npts(f));
        regs.push_back(I);
    }
    if (regs.empty()) return MX_OK;
    if (list.empty()) list.push_back(cands[0]);  // keep the table non-empty
    Table tab(c);
    const size_t orr = tab.add(regs);
    const size_t oc = tab.add(list);
    int rc = tab.upload();
    if (rc) return rc;
    MX_LAUNCH(c, "k_fr_crse_init", k_fr_crse_init, grid_for(maxp, (int)regs.size(), 1, NTF),
              dim3(NTF), tab.dev<FrInit>(orr), tab.dev<FrCand>(oc), R.d, scomp, dcomp, ncomp, mult);
    return MX_OK;
}

int flux_reg_fine_add(mx_adv_ctx* c, int lev, const mx_multifab* flux, int dir, int scomp,
                      int dcomp, int ncomp, double mult) {
    mx_adv_ctx::FluxRegDev& R = c->freg[lev];
    const int dim = c->p.dim;
    int r[3];
    if (!fr_ratio(c, lev, r)) return MX_ERR_ARG;
    std::vector<FrAdd> regs;
    long long maxp = 0;
    for (size_t q = 0; q < R.box.size(); ++q) {
        if ((int)(q % (2 * dim)) % dim != dir) continue;
        const HBox f = hbox(R.box[q]);
        FrAdd A;
        A.off = R.off[q];
        set_lohi(A.lo, A.hi, f);
        A.fine = to_da(flux->fab[q / (2 * dim)]);
        maxp = std::max(maxp, npts(f));
        regs.push_back(A);
    }
    if (regs.empty()) return MX_OK;
    Table tab(c);
    const size_t orr = tab.add(regs);
    int rc = tab.upload();
    if (rc) return rc;
    MX_LAUNCH(c, "k_fr_fine_add", k_fr_fine_add, grid_for(maxp, (int)regs.size(), 1, NTF),
              dim3(NTF), tab.dev<FrAdd>(orr), R.d, dir, r[0], r[1], r[2], scomp, dcomp, ncomp,
              mult);
    return MX_OK;
}

// the register updates of MakeRhoXFlux (rhox: components start_comp ..
// start_comp + num_comp - 1 and Rho) or MakeRhoHFlux (RhoH) on level lev with
// reflux_type 2: CrseInit into the register of lev + 1 with
// -dt * area, FineAdd into the register of lev with dt * area, area the
// level's face area of each direction (MaestroMakeFlux.cpp:420-451,
// 1035-1058)
int flux_reg_ops(mx_adv_ctx* c, int lev, const mx_face_multifab* sflux, bool rhox,
                 int start_comp, int num_comp) {
    const mx_adv_params& p = c->p;
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
    int rc;
    const bool crse = lev + 1 <= p.finest_level && lev + 1 < MX_MAX_LEV && c->freg[lev + 1].d;
    const bool fine = lev >= 1 && c->freg[lev].d;
    const int need = rhox ? std::max(start_comp + num_comp, p.comp_rho + 1) : p.comp_rhoh + 1;
    if ((crse && c->freg[lev + 1].ncomp < need) || (fine && c->freg[lev].ncomp < need))
        return MX_ERR_ARG;
    for (int i = 0; i < dim && crse; ++i) {
        const double m = -1.0 * dt * area[i];
        if (rhox) {
            if ((rc = flux_reg_crse_init(c, lev + 1, &sflux->dir[i], i, start_comp, start_comp,
                                         num_comp, m)))
                return rc;
            if ((rc = flux_reg_crse_init(c, lev + 1, &sflux->dir[i], i, p.comp_rho, p.comp_rho, 1,
                                         m)))
                return rc;
        } else if ((rc = flux_reg_crse_init(c, lev + 1, &sflux->dir[i], i, p.comp_rhoh,
                                              p.comp_rhoh, 1, m))) {
            return rc;
        }
    }
    for (int i = 0; i < dim && fine; ++i) {
        const double m = 1.0 * dt * area[i];
        if (rhox) {
            if ((rc = flux_reg_fine_add(c, lev, &sflux->dir[i], i, start_comp, start_comp,
                                        num_comp, m)))
                return rc;
            if ((rc = flux_reg_fine_add(c, lev, &sflux->dir[i], i, p.comp_rho, p.comp_rho, 1, m)))
                return rc;
        } else if ((rc = flux_reg_fine_add(c, lev, &sflux->dir[i], i, p.comp_rhoh, p.comp_rhoh, 1,
                                             m))) {
            return rc;
        }
    }
    return MX_OK;
}

}  // namespace mxh

extern "C" {

int mx_adv_flux_reg_define(mx_adv_ctx* c, int lev, int ncomp) {
    return flux_reg_define(c, lev, ncomp);
}

int mx_adv_flux_reg_setval(mx_adv_ctx* c, int lev, double v) {
    int rc = fr_check(c, lev);
    if (rc) return rc;
    const long long n = c->freg[lev].total;
    MX_LAUNCH(c, "k_fr_fill", k_fr_fill, dim3((unsigned)((n + NTF - 1) / NTF)), dim3(NTF),
              c->freg[lev].d, n, v);
    return MX_OK;
}

int mx_adv_flux_reg_crse_init(mx_adv_ctx* c, int lev, const mx_multifab* flux, int dir,
                              int srccomp, int destcomp, int numcomp, double mult) {
    int rc = fr_check(c, lev);
    if (rc) return rc;
    if (dir < 0 || dir >= c->p.dim || srccomp < 0 || destcomp < 0 || numcomp < 1 ||
        destcomp + numcomp > c->freg[lev].ncomp)
        return MX_ERR_ARG;
    if ((rc = check_dir_flux(c, lev - 1, flux, dir, srccomp + numcomp))) return rc;
    return flux_reg_crse_init(c, lev, flux, dir, srccomp, destcomp, numcomp, mult);
}

int mx_adv_flux_reg_fine_add(mx_adv_ctx* c, int lev, const mx_multifab* flux, int dir,
                             int srccomp, int destcomp, int numcomp, double mult) {
    int rc = fr_check(c, lev);
    if (rc) return rc;
    if (dir < 0 || dir >= c->p.dim || srccomp < 0 || destcomp < 0 || numcomp < 1 ||
        destcomp + numcomp > c->freg[lev].ncomp)
        return MX_ERR_ARG;
    if ((rc = check_dir_flux(c, lev, flux, dir, srccomp + numcomp))) return rc;
    return flux_reg_fine_add(c, lev, flux, dir, srccomp, destcomp, numcomp, mult);
}

int mx_adv_flux_reg_reflux(mx_adv_ctx* c, int lev, mx_multifab* state, double scale, int scomp,
                           int dcomp, int ncomp) {
    int rc = fr_check(c, lev);
    if (rc) return rc;
    const mx_adv_ctx::FluxRegDev& R = c->freg[lev];
    if (scomp < 0 || dcomp < 0 || ncomp < 1 || scomp + ncomp > R.ncomp) return MX_ERR_ARG;
    if ((rc = check_mf(c, lev - 1, state, dcomp + ncomp, 0, -1))) return rc;
    const int dim = c->p.dim;
    const double* dx = c->p.dx[lev - 1];
    const double vol = dim == 3 ? dx[0] * dx[1] * dx[2] : dx[0] * dx[1];
    // periodic shifts of the coarse domain, 0 included
    std::vector<std::array<int, 3>> shifts;
    for (int a = -1; a <= 1; ++a)
        for (int b = -1; b <= 1; ++b)
            for (int e = -1; e <= 1; ++e) {
                const int s[3] = {a, b, e};
                bool ok = true;
                std::array<int, 3> sh{0, 0, 0};
                for (int d = 0; d < 3; ++d) {
                    if (s[d] != 0 && (d >= dim || !c->p.is_periodic[d])) ok = false;
                    sh[d] = s[d] * (c->p.domain_hi[lev - 1][d] - c->p.domain_lo[lev - 1][d] + 1);
                }
                if (ok) shifts.push_back(sh);
            }
    const auto& cg = c->grids[lev - 1];
    std::vector<FrCell> cells;
    std::vector<FrRef> refs;
    long long maxp = 0;
    for (size_t b = 0; b < cg.size(); ++b) {
        const HBox cb = hbox(cg[b]);
        FrCell C;
        C.s = to_da(state->fab[b]);
        set_lohi(C.lo, C.hi, cb);
        // faces a cell of the box can read: the box's faces in every direction
        HBox near = cb;
        for (int d = 0; d < dim; ++d) near.hi[d] += 1;
        for (int o = 0; o < 2 * dim; ++o) {
            C.ref_off[o] = (int)refs.size();
            for (size_t q = o; q < R.box.size(); q += 2 * dim)
                for (const auto& sh : shifts) {
                    HBox f = hbox(R.box[q]);
                    for (int d = 0; d < 3; ++d) {
                        f.lo[d] += sh[d];
                        f.hi[d] += sh[d];
                    }
                    if (!hit(f, near)) continue;
                    FrRef F;
                    F.off = R.off[q];
                    set_lohi(F.lo, F.hi, f);
                    F.pad_ = 0;
                    refs.push_back(F);
                }
        }
        for (int o = 2 * dim; o < 7; ++o) C.ref_off[o] = (int)refs.size();
        maxp = std::max(maxp, npts(cb));
        cells.push_back(C);
    }
    if (refs.empty()) refs.push_back(FrRef{0, {0, 0, 0}, {-1, -1, -1}, 0});
    Table tab(c);
    const size_t oc = tab.add(cells);
    const size_t orf = tab.add(refs);
    if ((rc = tab.upload())) return rc;
    MX_LAUNCH(c, "k_fr_reflux", k_fr_reflux, grid_for(maxp, (int)cells.size(), 1, NTF), dim3(NTF),
              tab.dev<FrCell>(oc), tab.dev<FrRef>(orf), R.d, dim, scale, vol, scomp, dcomp, ncomp);
    return MX_OK;
}

int mx_adv_flux_reg_read(mx_adv_ctx* c, int lev, double* host, long long cap, long long* n) {
    int rc = fr_check(c, lev);
    if (rc) return rc;
    if (!n) return MX_ERR_ARG;
    *n = c->freg[lev].total;
    if (!host) return MX_OK;
    if (cap < c->freg[lev].total) return MX_ERR_ARG;
    MX_HIP_CHECK(hipStreamSynchronize(c->stream));
    MX_HIP_CHECK(hipMemcpy(host, c->freg[lev].d, (size_t)c->freg[lev].total * sizeof(double),
                           hipMemcpyDeviceToHost));
    return MX_OK;
}

}  // extern "C"
