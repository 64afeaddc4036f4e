// mx_update.hip -- UpdateScal and UpdateVel: the flux-divergence update that
// closes the advection step, on the valid cells of one level.
//
// One thread per cell, every box of the level in one launch (box tables as
// in the other stages).  The arithmetic keeps the reference's operation
// order, so the results equal the CPU path's bit for bit.  The species pass
// (density as the sum of the species updates, auxiliary rescaling, density
// floor, redistribution of negative species) runs per cell in one thread; for
// up to MAXS species it keeps them in registers, beyond that it works on the
// output array as the reference does.  Both forms evaluate the same
// operations in the same order.
//
// Both stages stream their inputs once: HBM-bound.  Algorithmic bytes per
// cell (3-D, each face counted once): UpdateScal species 8 (2 + 6 NumSpec)
// (old and new density and species, three fluxes and the force of each
// species), enthalpy 8 x 6, UpdateVel 8 x 22 planar (old and new velocity,
// force, MAC velocity, nine edge velocities, w0; +1 with the sponge).
#include "mx_internal.H"

using namespace mxd;
using namespace mxh;

namespace {

constexpr int NTU = 256;
constexpr int MAXS = 8;  // species held in registers

struct UsBox {
    int lo[3], hi[3];
    DA so, sn, fo, sf[3];
};

struct UsArgs {
    double dt, dx[3];
    double half_cutoff;  // 0.5 * base_cutoff_density
    int dim;
    int Rho, RhoH, FirstSpec, NumSpec, FirstAux, NumAux;
};

__device__ __forceinline__ bool cell(const UsBox& B, int& i, int& j, int& k) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    return region_point(idx, B.lo, B.hi, i, j, k);
}

// the flux divergence of component comp, accumulated as the species loop of
// UpdateScal does (MaestroAdvection.cpp:170-180)
__device__ __forceinline__ double div_spec(const UsBox& B, const UsArgs& A, int i, int j, int k,
                                           int comp) {
    double divterm = (B.sf[0](i + 1, j, k, comp) - B.sf[0](i, j, k, comp)) / A.dx[0];
    divterm += (B.sf[1](i, j + 1, k, comp) - B.sf[1](i, j, k, comp)) / A.dx[1];
    if (A.dim == 3) divterm += (B.sf[2](i, j, k + 1, comp) - B.sf[2](i, j, k, comp)) / A.dx[2];
    return divterm;
}

// species, density, auxiliaries, floor and negative species of one cell
// (MaestroAdvection.cpp:164-249).  REG: the species live in x[0..MAXS)
template <bool REG>
__global__ void __launch_bounds__(NTU) k_update_rhox(const UsBox* __restrict__ boxes, UsArgs A) {
    const UsBox& B = boxes[blockIdx.y];
    int i, j, k;
    if (!cell(B, i, j, k)) return;
    const int S0 = A.FirstSpec, NS = A.NumSpec;
    double x[MAXS], xo[MAXS];
    if (REG) {
#pragma unroll
        for (int n = 0; n < MAXS; ++n) {
            x[n] = 0.0;
            xo[n] = 0.0;
            if (n < NS) {
                const int comp = S0 + n;
                const double divterm = div_spec(B, A, i, j, k, comp);
                xo[n] = B.so(i, j, k, comp);
                x[n] = xo[n] + A.dt * (-divterm + B.fo(i, j, k, comp));
            }
        }
    } else {
        for (int n = 0; n < NS; ++n) {
            const int comp = S0 + n;
            const double divterm = div_spec(B, A, i, j, k, comp);
            B.sn(i, j, k, comp) = B.so(i, j, k, comp) + A.dt * (-divterm + B.fo(i, j, k, comp));
        }
    }
    // density: the sum of the species updates
    const double rho_old = B.so(i, j, k, A.Rho);
    double rho = rho_old;
    bool neg = false;
    if (REG) {
#pragma unroll
        for (int n = 0; n < MAXS; ++n) {
            if (n < NS) {
                rho += x[n] - xo[n];
                if (x[n] < 0.0) neg = true;
            }
        }
    } else {
        for (int n = 0; n < NS; ++n) {
            const double v = B.sn(i, j, k, S0 + n);
            rho += v - B.so(i, j, k, S0 + n);
            if (v < 0.0) neg = true;
        }
    }
    for (int comp = A.FirstAux; comp < A.FirstAux + A.NumAux; ++comp)
        B.sn(i, j, k, comp) = B.so(i, j, k, comp) * rho / rho_old;
    // density floor
    if (rho < A.half_cutoff) {
        const double s = A.half_cutoff / rho;
        if (REG) {
#pragma unroll
            for (int n = 0; n < MAXS; ++n)
                if (n < NS) x[n] *= s;
        } else {
            for (int n = 0; n < NS; ++n) B.sn(i, j, k, S0 + n) *= s;
        }
        rho = A.half_cutoff;
    }
    // no species may leave negative: each negative one takes its deficit from
    // the non-negative others in proportion, in component order
    if (neg) {
        if (REG) {
#pragma unroll
            for (int c1 = 0; c1 < MAXS; ++c1) {
                if (c1 < NS && x[c1] < 0.0) {
                    const double delta = -x[c1];
                    double sumX = 0.0;
#pragma unroll
                    for (int c2 = 0; c2 < MAXS; ++c2)
                        if (c2 < NS && c2 != c1 && x[c2] >= 0.0) sumX += x[c2];
#pragma unroll
                    for (int c2 = 0; c2 < MAXS; ++c2) {
                        if (c2 < NS && c2 != c1 && x[c2] >= 0.0) {
                            const double frac = x[c2] / sumX;
                            x[c2] -= frac * delta;
                        }
                    }
                    x[c1] = 0.0;
                }
            }
        } else {
            for (int c1 = 0; c1 < NS; ++c1) {
                if (B.sn(i, j, k, S0 + c1) < 0.0) {
                    const double delta = -B.sn(i, j, k, S0 + c1);
                    double sumX = 0.0;
                    for (int c2 = 0; c2 < NS; ++c2) {
                        const double v = B.sn(i, j, k, S0 + c2);
                        if (c2 != c1 && v >= 0.0) sumX += v;
                    }
                    for (int c2 = 0; c2 < NS; ++c2) {
                        const double v = B.sn(i, j, k, S0 + c2);
                        if (c2 != c1 && v >= 0.0) {
                            const double frac = v / sumX;
                            B.sn(i, j, k, S0 + c2) = v - frac * delta;
                        }
                    }
                    B.sn(i, j, k, S0 + c1) = 0.0;
                }
            }
        }
    }
    if (REG) {
#pragma unroll
        for (int n = 0; n < MAXS; ++n)
            if (n < NS) B.sn(i, j, k, S0 + n) = x[n];
    }
    B.sn(i, j, k, A.Rho) = rho;
}

// enthalpy (MaestroAdvection.cpp:120-136)
__global__ void __launch_bounds__(NTU) k_update_rhoh(const UsBox* __restrict__ boxes, UsArgs A) {
    const UsBox& B = boxes[blockIdx.y];
    int i, j, k;
    if (!cell(B, i, j, k)) return;
    const int c = A.RhoH;
    double divterm = (B.sf[0](i + 1, j, k, c) - B.sf[0](i, j, k, c)) / A.dx[0] +
                     (B.sf[1](i, j + 1, k, c) - B.sf[1](i, j, k, c)) / A.dx[1];
    if (A.dim == 3) divterm = divterm + (B.sf[2](i, j, k + 1, c) - B.sf[2](i, j, k, c)) / A.dx[2];
    B.sn(i, j, k, c) = B.so(i, j, k, c) + A.dt * (-divterm + B.fo(i, j, k, c));
}

struct UvBox {
    int lo[3], hi[3];
    DA uo, un, fo, sp, w0c;
    DA um[3], ue[3], w0m[3];
};

struct UvArgs {
    double dt, dx[3];
    int sponge;
};

// (Utilde . grad) Utilde of component n, as the reference sums it
template <int DIM>
__device__ __forceinline__ double ugrad(const UvBox& B, const UvArgs& A, int i, int j, int k,
                                        int n, double ubar, double vbar, double wbar) {
    double g = ubar * (B.ue[0](i + 1, j, k, n) - B.ue[0](i, j, k, n)) / A.dx[0] +
               vbar * (B.ue[1](i, j + 1, k, n) - B.ue[1](i, j, k, n)) / A.dx[1];
    if (DIM == 3) g = g + wbar * (B.ue[2](i, j, k + 1, n) - B.ue[2](i, j, k, n)) / A.dx[2];
    return g;
}

// UpdateVel, planar (MaestroAdvection.cpp:335-422) or spherical (:423-505)
template <int DIM, bool SPH>
__global__ void __launch_bounds__(NTU) k_update_vel(const UvBox* __restrict__ boxes, UvArgs A) {
    const UvBox& B = boxes[blockIdx.y];
    int i, j, k;
    {
        const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
        if (!region_point(idx, B.lo, B.hi, i, j, k)) return;
    }
    const double dt = A.dt;
    const double ubar = 0.5 * (B.um[0](i, j, k) + B.um[0](i + 1, j, k));
    const double vbar = 0.5 * (B.um[1](i, j, k) + B.um[1](i, j + 1, k));
    const double wbar = DIM == 3 ? 0.5 * (B.um[2](i, j, k) + B.um[2](i, j, k + 1)) : 0.0;
    double un[DIM];
#pragma unroll
    for (int n = 0; n < DIM; ++n)
        un[n] = B.uo(i, j, k, n) - dt * ugrad<DIM>(B, A, i, j, k, n, ubar, vbar, wbar) +
                dt * B.fo(i, j, k, n);
    if (!SPH) {
        // (w0 . grad) Utilde with w0 the cell-centred w0_cart of this cell
        // and the one above it
        const double w0bar =
            DIM == 2 ? 0.5 * (B.w0c(i, j, k, DIM - 1) + B.w0c(i, j + 1, k, DIM - 1))
                     : 0.5 * (B.w0c(i, j, k, DIM - 1) + B.w0c(i, j, k + 1, DIM - 1));
        const double sp = A.sponge ? B.sp(i, j, k) : 1.0;
#pragma unroll
        for (int n = 0; n < DIM; ++n) {
            if (DIM == 2)
                un[n] -= dt * w0bar * (B.ue[1](i, j + 1, k, n) - B.ue[1](i, j, k, n)) / A.dx[1];
            else
                un[n] -= dt * w0bar * (B.ue[2](i, j, k + 1, n) - B.ue[2](i, j, k, n)) / A.dx[2];
            if (A.sponge) un[n] *= sp;
        }
    } else {
        const double wx = B.w0m[0](i, j, k) + B.w0m[0](i + 1, j, k);
        const double wy = B.w0m[1](i, j, k) + B.w0m[1](i, j + 1, k);
        const double wz = B.w0m[2](i, j, k) + B.w0m[2](i, j, k + 1);
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            const double gx = (B.ue[0](i + 1, j, k, n) - B.ue[0](i, j, k, n)) / A.dx[0];
            const double gy = (B.ue[1](i, j + 1, k, n) - B.ue[1](i, j, k, n)) / A.dx[1];
            const double gz = (B.ue[2](i, j, k + 1, n) - B.ue[2](i, j, k, n)) / A.dx[2];
            const double w0g = gx * 0.5 * wx + gy * 0.5 * wy + gz * 0.5 * wz;
            un[n] -= dt * w0g;
        }
        if (A.sponge) {
            const double sp = B.sp(i, j, k);
#pragma unroll
            for (int n = 0; n < 3; ++n) un[n] *= sp;
        }
    }
#pragma unroll
    for (int n = 0; n < DIM; ++n) B.un(i, j, k, n) = un[n];
}

long long box_points(const mx_box& b) {
    long long np = 1;
    for (int d = 0; d < 3; ++d) np *= (long long)(b.hi[d] - b.lo[d] + 1);
    return np;
}

}  // namespace

namespace mxh {

int update_scal(mx_adv_ctx* c, int lev, const mx_multifab* stateold, mx_multifab* statenew,
                const mx_face_multifab* sflux, const mx_multifab* force, int start_comp,
                double base_cutoff_density) {
    const mx_adv_params& p = c->p;
    const auto& grids = c->grids[lev];
    std::vector<UsBox> boxes;
    long long maxpts = 0;
    for (size_t b = 0; b < grids.size(); ++b) {
        UsBox B;
        HBox g = hbox(grids[b]);
        for (int d = 0; d < 3; ++d) {
            B.lo[d] = g.lo[d];
            B.hi[d] = g.hi[d];
        }
        B.so = to_da(stateold->fab[b]);
        B.sn = to_da(statenew->fab[b]);
        B.fo = to_da(force->fab[b]);
        for (int d = 0; d < 3; ++d)
            B.sf[d] = d < p.dim ? to_da(sflux->dir[d].fab[b]) : null_da();
        maxpts = std::max(maxpts, box_points(grids[b]));
        boxes.push_back(B);
    }
    UsArgs A;
    A.dt = p.dt;
    for (int d = 0; d < 3; ++d) A.dx[d] = p.dx[lev][d];
    A.half_cutoff = 0.5 * base_cutoff_density;
    A.dim = p.dim;
    A.Rho = p.comp_rho;
    A.RhoH = p.comp_rhoh;
    A.FirstSpec = p.first_spec;
    A.NumSpec = p.num_spec;
    A.FirstAux = p.first_spec + p.num_spec;
    A.NumAux = p.nscal - A.FirstAux;
    Table tab(c);
    size_t ob = tab.add(boxes);
    int rc = tab.upload();
    if (rc) return rc;
    const dim3 gr = grid_for(maxpts, (int)boxes.size(), 1, NTU);
    if (start_comp == p.comp_rhoh) {
        MX_LAUNCH(c, "k_update_rhoh", k_update_rhoh, gr, dim3(NTU), tab.dev<UsBox>(ob), A);
    } else if (p.num_spec <= MAXS) {
        MX_LAUNCH(c, "k_update_rhox", k_update_rhox<true>, gr, dim3(NTU), tab.dev<UsBox>(ob), A);
    } else {
        MX_LAUNCH(c, "k_update_rhox", k_update_rhox<false>, gr, dim3(NTU), tab.dev<UsBox>(ob), A);
    }
    return MX_OK;
}

int update_vel(mx_adv_ctx* c, int lev, const mx_multifab* uold, mx_multifab* unew,
               const mx_face_multifab* umac, const mx_face_multifab* uedge,
               const mx_multifab* force, const mx_multifab* sponge,
               const mx_face_multifab* w0mac, const mx_multifab* w0_cart, int do_sponge) {
    const mx_adv_params& p = c->p;
    const auto& grids = c->grids[lev];
    const bool sph = p.dim == 3 && p.spherical;
    std::vector<UvBox> boxes;
    long long maxpts = 0;
    for (size_t b = 0; b < grids.size(); ++b) {
        UvBox B;
        HBox g = hbox(grids[b]);
        for (int d = 0; d < 3; ++d) {
            B.lo[d] = g.lo[d];
            B.hi[d] = g.hi[d];
        }
        B.uo = to_da(uold->fab[b]);
        B.un = to_da(unew->fab[b]);
        B.fo = to_da(force->fab[b]);
        B.sp = do_sponge ? to_da(sponge->fab[b]) : null_da();
        B.w0c = sph ? null_da() : to_da(w0_cart->fab[b]);
        for (int d = 0; d < 3; ++d) {
            B.um[d] = d < p.dim ? to_da(umac->dir[d].fab[b]) : null_da();
            B.ue[d] = d < p.dim ? to_da(uedge->dir[d].fab[b]) : null_da();
            B.w0m[d] = sph ? to_da(w0mac->dir[d].fab[b]) : null_da();
        }
        maxpts = std::max(maxpts, box_points(grids[b]));
        boxes.push_back(B);
    }
    UvArgs A;
    A.dt = p.dt;
    for (int d = 0; d < 3; ++d) A.dx[d] = p.dx[lev][d];
    A.sponge = do_sponge ? 1 : 0;
    Table tab(c);
    size_t ob = tab.add(boxes);
    int rc = tab.upload();
    if (rc) return rc;
    const dim3 gr = grid_for(maxpts, (int)boxes.size(), 1, NTU);
    if (p.dim == 2)
        MX_LAUNCH(c, "k_update_vel", (k_update_vel<2, false>), gr, dim3(NTU), tab.dev<UvBox>(ob), A);
    else if (sph)
        MX_LAUNCH(c, "k_update_vel", (k_update_vel<3, true>), gr, dim3(NTU), tab.dev<UvBox>(ob), A);
    else
        MX_LAUNCH(c, "k_update_vel", (k_update_vel<3, false>), gr, dim3(NTU), tab.dev<UvBox>(ob), A);
    return MX_OK;
}

}  // namespace mxh

// ---------------------------------------------------------------------------
// C-ABI

extern "C" {

int mx_adv_update_scal(mx_adv_ctx* c, int lev, const mx_multifab* stateold,
                       mx_multifab* statenew, const mx_face_multifab* sflux,
                       const mx_multifab* force, int start_comp, int num_comp,
                       const mx_multifab* p0_cart, double base_cutoff_density,
                       int do_eos_h_above_cutoff) {
    (void)p0_cart;
    int rc = check_level(c, lev);
    if (rc) return rc;
    const mx_adv_params& p = c->p;
    if (do_eos_h_above_cutoff) return MX_ERR_UNSUPPORTED;
    if (start_comp == p.comp_rhoh) {
        if (num_comp != 1) return MX_ERR_ARG;
    } else if (start_comp == p.first_spec) {
        if (num_comp != p.num_spec) return MX_ERR_ARG;
    } else {
        return MX_ERR_ARG;
    }
    if ((rc = check_mf(c, lev, stateold, p.nscal, 0, -1))) return rc;
    if ((rc = check_mf(c, lev, statenew, p.nscal, 0, -1))) return rc;
    if ((rc = check_mf(c, lev, force, p.nscal, 0, -1))) return rc;
    if ((rc = check_face(c, lev, sflux, p.nscal, 0))) return rc;
    return update_scal(c, lev, stateold, statenew, sflux, force, start_comp, base_cutoff_density);
}

int mx_adv_update_vel(mx_adv_ctx* c, int lev, const mx_multifab* uold, mx_multifab* unew,
                      const mx_face_multifab* umac, const mx_face_multifab* uedge,
                      const mx_multifab* force, const mx_multifab* sponge,
                      const mx_face_multifab* w0mac, const mx_multifab* w0_cart, int do_sponge) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    const mx_adv_params& p = c->p;
    const int dim = p.dim;
    if ((rc = check_mf(c, lev, uold, dim, 0, -1))) return rc;
    if ((rc = check_mf(c, lev, unew, dim, 0, -1))) return rc;
    if ((rc = check_mf(c, lev, force, dim, 0, -1))) return rc;
    if ((rc = check_face(c, lev, umac, 1, 0))) return rc;
    if ((rc = check_face(c, lev, uedge, dim, 0))) return rc;
    if (do_sponge && (rc = check_mf(c, lev, sponge, 1, 0, -1))) return rc;
    if (dim == 3 && p.spherical) {
        if ((rc = check_face(c, lev, w0mac, 1, 0))) return rc;
    } else {
        // w0_cart of the cell above each valid cell: one ghost cell
        if ((rc = check_mf(c, lev, w0_cart, dim, 1, -1))) return rc;
    }
    return update_vel(c, lev, uold, unew, umac, uedge, force, sponge, w0mac, w0_cart, do_sponge);
}

}  // extern "C"
