// mx_base.hip -- the base-state <-> Cartesian producers that feed the pass:
// Put1dArrayOnCart, MakeW0mac, MakeS0mac, PutInPertForm, ConvertRhoXToX and
// ConvertRhoHToH, one level each.
//
// One thread per cell (or face), every box of the level in one launch.  Each
// point evaluates the reference's formula for its own position with the
// reference's operation order, so the results equal the CPU path's bit for
// bit.  Spherical geometry is the non-exact base state: the radial index of
// a point is int(radius / dr_fine).  HBM-bound except for the spherical
// interpolation (a square root and a few divisions per point).
#include "mx_internal.H"

using namespace mxd;
using namespace mxh;

namespace {

constexpr int NTBS = 256;

struct BBox {
    int lo[3], hi[3];  // the point region (cells or faces)
    DA out, in;        // output; input (w0_cart / s0_cart / scal)
};

struct SphG {
    double center[3], prob_lo[3], dx[3];
    double drf;
    int nr_fine;
    int pad_ = 0;
    const double* rcc;
    const double* redge;
};

struct BaseA {
    const double* s0;  // row `lev` of the base state (spherical: row 0)
    int edge, vec, type, dim, outcomp, comp0, ncomp, flag, Rho;
    int pad_ = 0;
};

__device__ __forceinline__ double quad_interp(double x, double x0, double x1, double x2,
                                              double y0, double y1, double y2) {
    double y = y0 + (y1 - y0) / (x1 - x0) * (x - x0) +
               ((y2 - y1) / (x2 - x1) - (y1 - y0) / (x1 - x0)) / (x2 - x0) * (x - x0) * (x - x1);
    if (y > dmax(y0, dmax(y1, y2))) y = dmax(y0, dmax(y1, y2));
    if (y < dmin(y0, dmin(y1, y2))) y = dmin(y0, dmin(y1, y2));
    return y;
}

// the edge-centred base state at radius (w0_interp_type 1-3,
// MaestroFill3dData.cpp:211-253)
__device__ __forceinline__ double edge_value(const SphG& G, const double* s0, double radius,
                                             int type) {
    const double drf = G.drf;
    const int nr_fine = G.nr_fine;
    int index = int(radius / drf);
    const double rfac = (radius - double(index) * drf) / drf;
    double v = 0.0;
    if (type == 1) {
        v = rfac > 0.5 ? s0[index + 1] : s0[index];
    } else if (type == 2) {
        if (index < nr_fine)
            v = rfac * s0[index + 1] + (1.0 - rfac) * s0[index];
        else
            v = s0[nr_fine];
    } else if (type == 3) {
        if (index <= 0)
            index = 0;
        else if (index >= nr_fine - 1)
            index = nr_fine - 2;
        else if (radius - G.redge[index] < G.redge[index + 1])
            index--;
        v = quad_interp(radius, G.redge[index], G.redge[index + 1], G.redge[index + 2], s0[index],
                        s0[index + 1], s0[index + 2]);
    }
    return v;
}

// the bin-centred base state at radius (s0_interp_type 1-3,
// MaestroFill3dData.cpp:271-325; s0mac_interp_type 2-3, :1009-1188)
__device__ __forceinline__ double cell_value(const SphG& G, const double* s0, double radius,
                                             int type) {
    const double drf = G.drf;
    const int nr_fine = G.nr_fine;
    int index = int(radius / drf);
    double v = 0.0;
    if (type == 1) {
        v = s0[index];
    } else if (type == 2) {
        if (radius >= G.rcc[index]) {
            if (index >= nr_fine - 1)
                v = s0[nr_fine - 1];
            else
                v = s0[index + 1] * (radius - G.rcc[index]) / drf +
                    s0[index] * (G.rcc[index + 1] - radius) / drf;
        } else {
            if (index == 0)
                v = s0[index];
            else if (index > nr_fine - 1)
                v = s0[nr_fine - 1];
            else
                v = s0[index] * (radius - G.rcc[index - 1]) / drf +
                    s0[index - 1] * (G.rcc[index] - radius) / drf;
        }
    } else if (type == 3) {
        if (index == 0)
            index = 1;
        else if (index >= nr_fine - 1)
            index = nr_fine - 2;
        v = quad_interp(radius, G.rcc[index - 1], G.rcc[index], G.rcc[index + 1], s0[index - 1],
                        s0[index], s0[index + 1]);
    }
    return v;
}

// position of a point relative to the centre; h = 0.5 for a cell-centred
// coordinate, 0 for a face or node coordinate
__device__ __forceinline__ void pos(const SphG& G, int i, int j, int k, double h0, double h1,
                                    double h2, double& x, double& y, double& z) {
    x = G.prob_lo[0] + (double(i) + h0) * G.dx[0] - G.center[0];
    y = G.prob_lo[1] + (double(j) + h1) * G.dx[1] - G.center[1];
    z = G.prob_lo[2] + (double(k) + h2) * G.dx[2] - G.center[2];
}

__device__ __forceinline__ bool bpoint(const BBox& B, int& i, int& j, int& k) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    return region_point(idx, B.lo, B.hi, i, j, k);
}

// the value Put1dArrayOnCart gives cell (i,j,k) (MaestroFill3dData.cpp:49-339)
template <bool SPH>
__device__ __forceinline__ double put1d_scalar(const SphG& G, const BaseA& A, int i, int j,
                                               int k, double& x, double& y, double& z,
                                               double& radius) {
    if (!SPH) {
        const int r = A.dim == 2 ? j : k;
        return A.edge ? 0.5 * (A.s0[r] + A.s0[r + 1]) : A.s0[r];
    }
    pos(G, i, j, k, 0.5, 0.5, 0.5, x, y, z);
    radius = sqrt(x * x + y * y + z * z);
    return A.edge ? edge_value(G, A.s0, radius, A.type) : cell_value(G, A.s0, radius, A.type);
}

template <bool SPH>
__global__ void __launch_bounds__(NTBS) k_put1d(const BBox* __restrict__ boxes, SphG G, BaseA A) {
    const BBox& B = boxes[blockIdx.y];
    int i, j, k;
    if (!bpoint(B, i, j, k)) return;
    double x = 0.0, y = 0.0, z = 0.0, radius = 0.0;
    const double v = put1d_scalar<SPH>(G, A, i, j, k, x, y, z, radius);
    if (SPH && A.vec) {
        B.out(i, j, k, 0) = v * x / radius;
        B.out(i, j, k, 1) = v * y / radius;
        B.out(i, j, k, 2) = v * z / radius;
    } else {
        B.out(i, j, k, A.outcomp) = v;
    }
}

// PutInPertForm: scal[comp] -+ the cell value of s0 (MaestroConvert.cpp:36-62)
template <bool SPH>
__global__ void __launch_bounds__(NTBS) k_pert_form(const BBox* __restrict__ boxes, SphG G,
                                                    BaseA A) {
    const BBox& B = boxes[blockIdx.y];
    int i, j, k;
    if (!bpoint(B, i, j, k)) return;
    double x = 0.0, y = 0.0, z = 0.0, radius = 0.0;
    const double v = put1d_scalar<SPH>(G, A, i, j, k, x, y, z, radius);
    double& s = B.out(i, j, k, A.comp0);
    if (A.flag)
        s -= v;
    else
        s += v;
}

// ConvertRhoXToX / ConvertRhoHToH (MaestroConvert.cpp:64-114)
__global__ void __launch_bounds__(NTBS) k_convert(const BBox* __restrict__ boxes, BaseA A) {
    const BBox& B = boxes[blockIdx.y];
    int i, j, k;
    if (!bpoint(B, i, j, k)) return;
    const double rho = B.out(i, j, k, A.Rho);
    for (int c = A.comp0; c < A.comp0 + A.ncomp; ++c) {
        double& s = B.out(i, j, k, c);
        if (A.flag)
            s /= rho;
        else
            s *= rho;
    }
}

// w0 at node (i,j,k), component n (MakeW0mac type 4, :496-526)
__device__ __forceinline__ double node_w0(const SphG& G, const double* w0, int i, int j, int k,
                                          int n) {
    double x, y, z;
    pos(G, i, j, k, 0.0, 0.0, 0.0, x, y, z);
    const double radius = sqrt(x * x + y * y + z * z);
    const int index = int(radius / G.drf);
    const double rfac = (radius - double(index) * G.drf) / G.drf;
    double v;
    if (index < G.nr_fine)
        v = rfac * w0[index + 1] + (1.0 - rfac) * w0[index];
    else
        v = w0[G.nr_fine];
    if (radius == 0.0) return v;
    const double q = n == 0 ? x : (n == 1 ? y : z);
    return v * q / radius;
}

// MakeW0mac (MaestroFill3dData.cpp:464-710) and MakeS0mac (:712-1188): the
// faces normal to D of every box grown by one cell; S0: s0mac, else w0mac
template <int D, bool S0>
__global__ void __launch_bounds__(NTBS) k_face_base(const BBox* __restrict__ boxes, SphG G,
                                                    BaseA A) {
    const BBox& B = boxes[blockIdx.y];
    int i, j, k;
    if (!bpoint(B, i, j, k)) return;
    const int type = A.type;
    double v;
    if (type == 1) {
        // the average of the two cells beside the face
        const double a = S0 ? B.in(i - (D == 0), j - (D == 1), k - (D == 2))
                            : B.in(i - (D == 0), j - (D == 1), k - (D == 2), D);
        const double b = S0 ? B.in(i, j, k) : B.in(i, j, k, D);
        v = 0.5 * (a + b);
    } else if (!S0 && type == 4) {
        constexpr int a0 = D == 0 ? 1 : 0;  // the face's transverse directions
        constexpr int a1 = D == 2 ? 1 : 2;
        const int i1 = i + (a0 == 0), j1 = j + (a0 == 1), k1 = k + (a0 == 2);
        const int i2 = i + (a1 == 0), j2 = j + (a1 == 1), k2 = k + (a1 == 2);
        const int i3 = i1 + (a1 == 0), j3 = j1 + (a1 == 1), k3 = k1 + (a1 == 2);
        v = 0.25 * (node_w0(G, A.s0, i, j, k, D) + node_w0(G, A.s0, i1, j1, k1, D) +
                    node_w0(G, A.s0, i2, j2, k2, D) + node_w0(G, A.s0, i3, j3, k3, D));
    } else {
        double x, y, z;
        pos(G, i, j, k, D == 0 ? 0.0 : 0.5, D == 1 ? 0.0 : 0.5, D == 2 ? 0.0 : 0.5, x, y, z);
        const double radius = sqrt(x * x + y * y + z * z);
        if (S0) {
            v = cell_value(G, A.s0, radius, type);
        } else {
            const double q = D == 0 ? x : (D == 1 ? y : z);
            v = edge_value(G, A.s0, radius, type) * q / radius;
        }
    }
    B.out(i, j, k) = v;
}

// ---------------------------------------------------------------- forces
// MakeVelForce (MaestroForce.cpp:6-258) and ModifyScalForce (:261-466).  The
// Cartesian base-state arrays the reference builds with Put1dArrayOnCart
// (rho0, grav_cell, gradw0, divw0, s0_edge) are evaluated per cell by the
// same formulas.

struct VfBox {
    int lo[3], hi[3];
    DA vf, rho, gpi, wf, w0c, nrm;
    DA ue[3];
};

struct VfArgs {
    const double* rho0;    // row lev (planar) or 0 (spherical)
    const double* grav;
    const double* gradw0;  // spherical: nr_fine values
    double cutoff;         // buoyancy_cutoff_factor * base_cutoff_density
    double dx[3];
    int dim, domhi, add_utilde, s0_type;
};

// the cell value of a bin-centred base state (Put1dArrayOnCart, scalar)
template <bool SPH>
__device__ __forceinline__ double cell_base(const SphG& G, const double* s0, int type, int dim,
                                            int i, int j, int k, double radius) {
    if (!SPH) return s0[dim == 2 ? j : k];
    return cell_value(G, s0, radius, type);
}

template <bool SPH>
__global__ void __launch_bounds__(NTBS) k_vel_force(const VfBox* __restrict__ boxes, SphG G,
                                                    VfArgs A) {
    const VfBox& B = boxes[blockIdx.y];
    int i, j, k;
    {
        const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
        if (!region_point(idx, B.lo, B.hi, i, j, k)) return;
    }
    const int dim = A.dim;
    double x = 0.0, y = 0.0, z = 0.0, radius = 0.0;
    if (SPH) {
        pos(G, i, j, k, 0.5, 0.5, 0.5, x, y, z);
        radius = sqrt(x * x + y * y + z * z);
    }
    const double ro = B.rho(i, j, k);
    double rhopert = ro - cell_base<SPH>(G, A.rho0, A.s0_type, dim, i, j, k, radius);
    if (ro < A.cutoff) rhopert = 0.0;
    const double g = cell_base<SPH>(G, A.grav, A.s0_type, dim, i, j, k, radius);
    if (!SPH) {
        for (int d = 0; d < dim - 1; ++d) B.vf(i, j, k, d) = -B.gpi(i, j, k, d) / ro;
        double v = (rhopert * g - B.gpi(i, j, k, dim - 1)) / ro - B.wf(i, j, k, dim - 1);
        if (A.add_utilde) {
            if (dim == 2) {
                if (!(j <= -1 || j >= A.domhi))
                    v -= (B.ue[1](i, j + 1, k) + B.ue[1](i, j, k)) *
                         (B.w0c(i, j + 1, k, 1) - B.w0c(i, j, k, 1)) / (2.0 * A.dx[1]);
            } else {
                if (!(k <= -1 || k >= A.domhi))
                    v -= (B.ue[2](i, j, k + 1) + B.ue[2](i, j, k)) *
                         (B.w0c(i, j, k + 1, 2) - B.w0c(i, j, k, 2)) / (2.0 * A.dx[2]);
            }
        }
        B.vf(i, j, k, dim - 1) = v;
    } else {
        const double gc[3] = {g * x / radius, g * y / radius, g * z / radius};
        double v[3];
#pragma unroll
        for (int d = 0; d < 3; ++d)
            v[d] = (rhopert * gc[d] - B.gpi(i, j, k, d)) / ro - B.wf(i, j, k, d);
        if (A.add_utilde == 1) {
            const double gw = cell_value(G, A.gradw0, radius, A.s0_type);
            const double ut = 0.5 * (B.ue[0](i, j, k) + B.ue[0](i + 1, j, k)) * B.nrm(i, j, k, 0) +
                              0.5 * (B.ue[1](i, j, k) + B.ue[1](i, j + 1, k)) * B.nrm(i, j, k, 1) +
                              0.5 * (B.ue[2](i, j, k) + B.ue[2](i, j, k + 1)) * B.nrm(i, j, k, 2);
#pragma unroll
            for (int d = 0; d < 3; ++d) v[d] -= ut * gw * B.nrm(i, j, k, d);
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) B.vf(i, j, k, d) = v[d];
    }
}

struct SfBox {
    int lo[3], hi[3];
    DA f, s, s0c, w0c;
    DA um[3];
};

struct SfArgs {
    const double* s0e;    // planar: row lev of s0_edge
    const double* divw0;  // spherical: nr_fine values
    double dx[3];
    int comp, fullform, dim, s0_type;
    int vlo, vhi, vper;
    int domlo[3], domhi[3];
};

template <bool SPH>
__global__ void __launch_bounds__(NTBS) k_scal_force(const SfBox* __restrict__ boxes, SphG G,
                                                     SfArgs A) {
    const SfBox& B = boxes[blockIdx.y];
    int i, j, k;
    {
        const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
        if (!region_point(idx, B.lo, B.hi, i, j, k)) return;
    }
    const int comp = A.comp;
    double& f = B.f(i, j, k, comp);
    const double sc = B.s(i, j, k, comp);
    const double* dx = A.dx;
    if (SPH) {
        double x, y, z;
        pos(G, i, j, k, 0.5, 0.5, 0.5, x, y, z);
        const double radius = sqrt(x * x + y * y + z * z);
        const double dv = cell_value(G, A.divw0, radius, A.s0_type);
        const double divumac = (B.um[0](i + 1, j, k) - B.um[0](i, j, k)) / dx[0] +
                               (B.um[1](i, j + 1, k) - B.um[1](i, j, k)) / dx[1] +
                               (B.um[2](i, j, k + 1) - B.um[2](i, j, k)) / dx[2];
        if (A.fullform) {
            f -= sc * (divumac + dv);
        } else {
            const double c0 = B.s0c(i, j, k);
            const double xhi = i < A.domhi[0] ? 0.5 * (c0 + B.s0c(i + 1, j, k)) : c0;
            const double xlo = i > A.domlo[0] ? 0.5 * (c0 + B.s0c(i - 1, j, k)) : c0;
            const double yhi = j < A.domhi[1] ? 0.5 * (c0 + B.s0c(i, j + 1, k)) : c0;
            const double ylo = j > A.domlo[1] ? 0.5 * (c0 + B.s0c(i, j - 1, k)) : c0;
            const double zhi = k < A.domhi[2] ? 0.5 * (c0 + B.s0c(i, j, k + 1)) : c0;
            const double zlo = k > A.domlo[2] ? 0.5 * (c0 + B.s0c(i, j, k - 1)) : c0;
            const double divs0u =
                (B.um[0](i + 1, j, k) * xhi - B.um[0](i, j, k) * xlo) / dx[0] +
                (B.um[1](i, j + 1, k) * yhi - B.um[1](i, j, k) * ylo) / dx[1] +
                (B.um[2](i, j, k + 1) * zhi - B.um[2](i, j, k) * zlo) / dx[2];
            f += -divs0u - (sc - c0) * (divumac + dv);
        }
    } else {
        const int dim = A.dim;
        double divu = (B.um[0](i + 1, j, k) - B.um[0](i, j, k)) / dx[0] +
                      (B.um[1](i, j + 1, k) - B.um[1](i, j, k)) / dx[1];
        if (dim == 3) divu = divu + (B.um[2](i, j, k + 1) - B.um[2](i, j, k)) / dx[2];
        if (dim == 2)
            divu += (B.w0c(i, j + 1, k, 1) - B.w0c(i, j, k, 1)) / dx[1];
        else
            divu += (B.w0c(i, j, k + 1, 2) - B.w0c(i, j, k, 2)) / dx[2];
        if (A.fullform) {
            f -= sc * divu;
        } else {
            // s0_edge put on the cells: the edge below each cell; above the
            // domain top, the top cell's value (first-order extrapolation) or
            // the periodic image
            const int r = dim == 2 ? j : k;
            int r1 = r + 1;
            if (r1 > A.vhi) r1 = A.vper ? r1 - (A.vhi - A.vlo + 1) : A.vhi;
            const double e0 = A.s0e[r], e1 = A.s0e[r1];
            const double c0 = B.s0c(i, j, k);
            double divs0u;
            if (dim == 2)
                divs0u = c0 * (B.um[0](i + 1, j, k) - B.um[0](i, j, k)) / dx[0] +
                         (B.um[1](i, j + 1, k) * e1 - B.um[1](i, j, k) * e0) / dx[1];
            else
                divs0u = c0 * ((B.um[0](i + 1, j, k) - B.um[0](i, j, k)) / dx[0] +
                               (B.um[1](i, j + 1, k) - B.um[1](i, j, k)) / dx[1]) +
                         (B.um[2](i, j, k + 1) * e1 - B.um[2](i, j, k) * e0) / dx[2];
            f += -(sc - c0) * divu - divs0u;
        }
    }
}

// the radial arrays of the spherical forces: gradw0 (MaestroForce.cpp:34-44)
// or divw0 (:277-291), nr_fine values
__global__ void __launch_bounds__(NTBS) k_radial_w0(double* out, const double* w0,
                                                    const double* redge, const double* rcc,
                                                    int nr, double drf, int what, int zero) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    double v = 0.0;
    if (!zero) {
        if (what == 0) {
            v = (w0[r + 1] - w0[r]) / drf;
        } else if (r < nr - 1) {
            const double dr_loc = redge[r + 1] - redge[r];
            v = (redge[r + 1] * redge[r + 1] * w0[r + 1] - redge[r] * redge[r] * w0[r]) /
                (dr_loc * rcc[r] * rcc[r]);
        }
    }
    out[r] = v;
}

SphG sph_g(const mx_adv_ctx* c, int lev, const mx_sph_geom* sg) {
    SphG G;
    for (int d = 0; d < 3; ++d) {
        G.center[d] = sg ? sg->center[d] : 0.0;
        G.prob_lo[d] = sg ? sg->prob_lo[d] : 0.0;
        G.dx[d] = c->p.dx[lev][d];
    }
    G.drf = sg ? sg->dr_fine : 0.0;
    G.nr_fine = sg ? sg->nr_fine : 0;
    G.rcc = sg ? sg->r_cc_loc : nullptr;
    G.redge = sg ? sg->r_edge_loc : nullptr;
    return G;
}

BaseA base_args(const mx_adv_ctx* c) {
    BaseA A;
    A.s0 = nullptr;
    A.edge = A.vec = A.type = 0;
    A.dim = c->p.dim;
    A.outcomp = 0;
    A.comp0 = A.ncomp = A.flag = 0;
    A.Rho = c->p.comp_rho;
    return A;
}

// the largest radial index int(radius / dr_fine) of any cell centre, face or
// node of box b grown by `grow` cells: the base state must reach it plus the
// interpolation stencil.  Evaluated at the corners of the region, where the
// distance from the centre is largest, with a margin of one bin for rounding.
int max_radial_index(const mx_adv_ctx* c, int lev, const mx_sph_geom* sg, const HBox& r) {
    double rmax = 0.0;
    for (int corner = 0; corner < 8; ++corner) {
        double s = 0.0;
        for (int d = 0; d < 3; ++d) {
            const int idx = (corner >> d) & 1 ? r.hi[d] + 1 : r.lo[d];
            const double x = sg->prob_lo[d] + double(idx) * c->p.dx[lev][d] - sg->center[d];
            s += x * x;
        }
        rmax = std::max(rmax, std::sqrt(s));
    }
    return (int)(rmax / sg->dr_fine) + 1;
}

int sph_check(const mx_adv_ctx* c, int lev, const mx_sph_geom* sg, const mx_basestate* s0,
              int need_extra, int grow_cells) {
    if (!sg || !sg->r_cc_loc || !sg->r_edge_loc || sg->nr_fine < 3 || !(sg->dr_fine > 0.0))
        return MX_ERR_ARG;
    if (c->p.use_exact_base_state) return MX_ERR_UNSUPPORTED;
    if (s0->nlev < 1) return MX_ERR_ARG;
    for (const mx_box& g : c->grids[lev]) {
        HBox r = grow(hbox(g), 3, grow_cells);
        const int imax = max_radial_index(c, lev, sg, r);
        // interpolation reads up to index + 2 of s0 and of r_edge_loc
        if (imax + need_extra >= s0->len || imax + need_extra > sg->nr_fine) return MX_ERR_ARG;
    }
    return MX_OK;
}

// upload the box table and launch: `launch(table, grid)` returns the launch
// status (MX_LAUNCH returns MX_ERR_HIP from it on a failed launch)
template <class F>
int launch_boxes(mx_adv_ctx* c, std::vector<BBox>& boxes, long long maxpts, F&& launch) {
    Table tab(c);
    size_t ob = tab.add(boxes);
    int rc = tab.upload();
    if (rc) return rc;
    return launch(tab.dev<BBox>(ob), grid_for(maxpts, (int)boxes.size(), 1, NTBS));
}

long long points(const HBox& r) {
    long long n = 1;
    for (int d = 0; d < 3; ++d) n *= (long long)(r.hi[d] - r.lo[d] + 1);
    return n;
}

}  // namespace

extern "C" {

int mx_adv_put_1d_array_on_cart(mx_adv_ctx* c, int lev, const mx_basestate* s0,
                                mx_multifab* s0_cart, int is_input_edge_centered,
                                int is_output_a_vector, const mx_sph_geom* sg) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    if (!s0 || !s0->p) return MX_ERR_ARG;
    const mx_adv_params& p = c->p;
    const bool sph = p.dim == 3 && p.spherical;
    const int nc = (is_output_a_vector && sph) ? 3 : (is_output_a_vector ? p.dim : 1);
    if ((rc = check_mf(c, lev, s0_cart, nc, 0, -1))) return rc;
    BaseA A = base_args(c);
    A.edge = is_input_edge_centered != 0;
    A.vec = is_output_a_vector != 0;
    A.outcomp = is_output_a_vector ? p.dim - 1 : 0;
    if (sph) {
        if ((rc = sph_check(c, lev, sg, s0, 2, 0))) return rc;
        A.type = A.edge ? sg->w0_interp_type : sg->s0_interp_type;
        if (A.type < 1 || A.type > 3) return MX_ERR_ARG;
        A.s0 = s0->p;
    } else {
        if (lev >= s0->nlev) return MX_ERR_ARG;
        if (s0->len < p.domain_hi[lev][p.dim - 1] + 1 + (A.edge ? 1 : 0)) return MX_ERR_ARG;
        A.s0 = s0->p + (long long)lev * s0->len;
    }
    const SphG G = sph_g(c, lev, sg);
    std::vector<BBox> boxes;
    long long maxpts = 0;
    for (size_t b = 0; b < c->grids[lev].size(); ++b) {
        BBox B;
        HBox r = hbox(c->grids[lev][b]);
        for (int d = 0; d < 3; ++d) {
            B.lo[d] = r.lo[d];
            B.hi[d] = r.hi[d];
        }
        B.out = to_da(s0_cart->fab[b]);
        B.in = null_da();
        maxpts = std::max(maxpts, points(r));
        boxes.push_back(B);
    }
    return launch_boxes(c, boxes, maxpts, [&](const BBox* db, dim3 gr) -> int {
        if (sph)
            MX_LAUNCH(c, "k_put1d", k_put1d<true>, gr, dim3(NTBS), db, G, A);
        else
            MX_LAUNCH(c, "k_put1d", k_put1d<false>, gr, dim3(NTBS), db, G, A);
        return MX_OK;
    });
}

static int face_base(mx_adv_ctx* c, int lev, const mx_basestate* s0, const mx_multifab* cart,
                     mx_face_multifab* mac, const mx_sph_geom* sg, bool is_s0) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    const mx_adv_params& p = c->p;
    if (p.dim != 3 || !p.spherical) return MX_ERR_UNSUPPORTED;
    if (!s0 || !s0->p || !sg) return MX_ERR_ARG;
    const int type = is_s0 ? sg->s0mac_interp_type : sg->w0mac_interp_type;
    if (type < 1 || type > (is_s0 ? 3 : 4)) return MX_ERR_ARG;
    // faces of each box grown by one cell: the node region reaches one cell
    // further than the faces
    if ((rc = sph_check(c, lev, sg, s0, 2, 2))) return rc;
    if ((rc = check_face(c, lev, mac, 1, 1))) return rc;
    if (type == 1 && (rc = check_mf(c, lev, cart, is_s0 ? 1 : 3, 2, -1))) return rc;
    BaseA A = base_args(c);
    A.type = type;
    A.s0 = s0->p;
    const SphG G = sph_g(c, lev, sg);
    for (int d = 0; d < 3; ++d) {
        std::vector<BBox> boxes;
        long long maxpts = 0;
        for (size_t b = 0; b < c->grids[lev].size(); ++b) {
            BBox B;
            HBox r = grow(nodal(hbox(c->grids[lev][b]), d), 3, 1);
            for (int q = 0; q < 3; ++q) {
                B.lo[q] = r.lo[q];
                B.hi[q] = r.hi[q];
            }
            B.out = to_da(mac->dir[d].fab[b]);
            B.in = type == 1 ? to_da(cart->fab[b]) : null_da();
            maxpts = std::max(maxpts, points(r));
            boxes.push_back(B);
        }
        rc = launch_boxes(c, boxes, maxpts, [&](const BBox* db, dim3 gr) -> int {
            if (is_s0) {
                if (d == 0) MX_LAUNCH(c, "k_s0mac", (k_face_base<0, true>), gr, dim3(NTBS), db, G, A);
                if (d == 1) MX_LAUNCH(c, "k_s0mac", (k_face_base<1, true>), gr, dim3(NTBS), db, G, A);
                if (d == 2) MX_LAUNCH(c, "k_s0mac", (k_face_base<2, true>), gr, dim3(NTBS), db, G, A);
            } else {
                if (d == 0) MX_LAUNCH(c, "k_w0mac", (k_face_base<0, false>), gr, dim3(NTBS), db, G, A);
                if (d == 1) MX_LAUNCH(c, "k_w0mac", (k_face_base<1, false>), gr, dim3(NTBS), db, G, A);
                if (d == 2) MX_LAUNCH(c, "k_w0mac", (k_face_base<2, false>), gr, dim3(NTBS), db, G, A);
            }
            return MX_OK;
        });
        if (rc) return rc;
    }
    return MX_OK;
}

int mx_adv_make_w0mac(mx_adv_ctx* c, int lev, const mx_basestate* w0, const mx_multifab* w0_cart,
                      mx_face_multifab* w0mac, const mx_sph_geom* sg) {
    return face_base(c, lev, w0, w0_cart, w0mac, sg, false);
}

int mx_adv_make_s0mac(mx_adv_ctx* c, int lev, const mx_basestate* s0, const mx_multifab* s0_cart,
                      mx_face_multifab* s0mac, const mx_sph_geom* sg) {
    return face_base(c, lev, s0, s0_cart, s0mac, sg, true);
}

static int cell_pass(mx_adv_ctx* c, int lev, mx_multifab* scal, const BaseA& A, bool pert,
                     bool sph, const SphG& G) {
    std::vector<BBox> boxes;
    long long maxpts = 0;
    for (size_t b = 0; b < c->grids[lev].size(); ++b) {
        BBox B;
        HBox r = hbox(c->grids[lev][b]);
        for (int d = 0; d < 3; ++d) {
            B.lo[d] = r.lo[d];
            B.hi[d] = r.hi[d];
        }
        B.out = to_da(scal->fab[b]);
        B.in = null_da();
        maxpts = std::max(maxpts, points(r));
        boxes.push_back(B);
    }
    return launch_boxes(c, boxes, maxpts, [&](const BBox* db, dim3 gr) -> int {
        if (!pert)
            MX_LAUNCH(c, "k_convert", k_convert, gr, dim3(NTBS), db, A);
        else if (sph)
            MX_LAUNCH(c, "k_pert_form", k_pert_form<true>, gr, dim3(NTBS), db, G, A);
        else
            MX_LAUNCH(c, "k_pert_form", k_pert_form<false>, gr, dim3(NTBS), db, G, A);
        return MX_OK;
    });
}

int mx_adv_put_in_pert_form(mx_adv_ctx* c, int lev, mx_multifab* scal, const mx_basestate* s0,
                            int comp, int flag, const mx_sph_geom* sg) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    const mx_adv_params& p = c->p;
    if (!s0 || !s0->p || comp < 0 || comp >= p.nscal) return MX_ERR_ARG;
    if ((rc = check_mf(c, lev, scal, p.nscal, 0, -1))) return rc;
    const bool sph = p.dim == 3 && p.spherical;
    BaseA A = base_args(c);
    A.comp0 = comp;
    A.flag = flag != 0;
    if (sph) {
        if ((rc = sph_check(c, lev, sg, s0, 2, 0))) return rc;
        A.type = sg->s0_interp_type;
        if (A.type < 1 || A.type > 3) return MX_ERR_ARG;
        A.s0 = s0->p;
    } else {
        if (lev >= s0->nlev || s0->len < p.domain_hi[lev][p.dim - 1] + 1) return MX_ERR_ARG;
        A.s0 = s0->p + (long long)lev * s0->len;
    }
    return cell_pass(c, lev, scal, A, true, sph, sph_g(c, lev, sg));
}

static int convert(mx_adv_ctx* c, int lev, mx_multifab* scal, int comp0, int ncomp, int flag) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    if ((rc = check_mf(c, lev, scal, c->p.nscal, 0, -1))) return rc;
    BaseA A = base_args(c);
    A.comp0 = comp0;
    A.ncomp = ncomp;
    A.flag = flag != 0;
    return cell_pass(c, lev, scal, A, false, false, sph_g(c, lev, nullptr));
}

// the spherical radial array `what` (0 gradw0, 1 divw0) in workspace
static int radial_w0(mx_adv_ctx* c, const mx_basestate* w0, const mx_sph_geom* sg, int what,
                     int zero, double** out) {
    const int nr = sg->nr_fine;
    int rc = ws_reserve(c, (size_t)nr * sizeof(double) + 4096);
    if (rc) return rc;
    *out = ws_take(c, (size_t)nr);
    MX_LAUNCH(c, "k_radial_w0", k_radial_w0, dim3((nr + NTBS - 1) / NTBS), dim3(NTBS), *out, w0->p,
              sg->r_edge_loc, sg->r_cc_loc, nr, sg->dr_fine, what, zero);
    return MX_OK;
}

int mx_adv_make_vel_force(mx_adv_ctx* c, int lev, mx_multifab* vel_force,
                          const mx_face_multifab* uedge, const mx_multifab* rho,
                          const mx_multifab* gpi, const mx_multifab* w0_force_cart,
                          const mx_multifab* w0_cart, const mx_multifab* normal,
                          const mx_basestate* rho0, const mx_basestate* grav_cell,
                          const mx_basestate* w0, double buoyancy_cutoff_factor,
                          double base_cutoff_density, int do_add_utilde_force,
                          int average_base_state, const mx_sph_geom* sg) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    const mx_adv_params& p = c->p;
    const int dim = p.dim;
    const bool sph = dim == 3 && p.spherical;
    if (!rho0 || !rho0->p || !grav_cell || !grav_cell->p) return MX_ERR_ARG;
    // the reference reads the flag as a truth value (planar) and as == 1
    // (spherical): only 0 and 1 mean the same to both
    if (do_add_utilde_force != 0 && do_add_utilde_force != 1) return MX_ERR_ARG;
    if ((rc = check_mf(c, lev, vel_force, dim, 0, -1))) return rc;
    if ((rc = check_mf(c, lev, rho, 1, 0, -1))) return rc;
    if ((rc = check_mf(c, lev, gpi, dim, 0, -1))) return rc;
    if ((rc = check_mf(c, lev, w0_force_cart, dim, 0, -1))) return rc;
    if (do_add_utilde_force && (rc = check_face(c, lev, uedge, 1, 0))) return rc;
    VfArgs A;
    A.cutoff = buoyancy_cutoff_factor * base_cutoff_density;
    for (int d = 0; d < 3; ++d) A.dx[d] = p.dx[lev][d];
    A.dim = dim;
    A.domhi = p.domain_hi[lev][dim - 1];
    A.add_utilde = do_add_utilde_force;
    A.s0_type = 0;
    A.gradw0 = nullptr;
    if (sph) {
        if (!w0 || !w0->p) return MX_ERR_ARG;
        if ((rc = sph_check(c, lev, sg, rho0, 2, 0))) return rc;
        if ((rc = sph_check(c, lev, sg, grav_cell, 2, 0))) return rc;
        if (w0->len < sg->nr_fine + 1) return MX_ERR_ARG;
        if ((rc = check_mf(c, lev, normal, 3, 0, -1))) return rc;
        A.s0_type = sg->s0_interp_type;
        if (A.s0_type < 1 || A.s0_type > 3) return MX_ERR_ARG;
        A.rho0 = rho0->p;
        A.grav = grav_cell->p;
        double* gw = nullptr;
        if (do_add_utilde_force == 1) {
            rc = radial_w0(c, w0, sg, 0, p.use_exact_base_state || average_base_state, &gw);
            if (rc) return rc;
        }
        A.gradw0 = gw;
    } else {
        const int need = p.domain_hi[lev][dim - 1] + 1;
        if (lev >= rho0->nlev || lev >= grav_cell->nlev || rho0->len < need ||
            grav_cell->len < need)
            return MX_ERR_ARG;
        A.rho0 = rho0->p + (long long)lev * rho0->len;
        A.grav = grav_cell->p + (long long)lev * grav_cell->len;
        if (do_add_utilde_force && (rc = check_mf(c, lev, w0_cart, dim, 1, -1))) return rc;
    }
    const SphG G = sph_g(c, lev, sg);
    std::vector<VfBox> boxes;
    long long maxpts = 0;
    for (size_t b = 0; b < c->grids[lev].size(); ++b) {
        VfBox B;
        HBox r = hbox(c->grids[lev][b]);
        for (int d = 0; d < 3; ++d) {
            B.lo[d] = r.lo[d];
            B.hi[d] = r.hi[d];
        }
        B.vf = to_da(vel_force->fab[b]);
        B.rho = to_da(rho->fab[b]);
        B.gpi = to_da(gpi->fab[b]);
        B.wf = to_da(w0_force_cart->fab[b]);
        B.w0c = (!sph && do_add_utilde_force) ? to_da(w0_cart->fab[b]) : null_da();
        B.nrm = sph ? to_da(normal->fab[b]) : null_da();
        for (int d = 0; d < 3; ++d)
            B.ue[d] = (do_add_utilde_force && d < dim) ? to_da(uedge->dir[d].fab[b]) : null_da();
        maxpts = std::max(maxpts, points(r));
        boxes.push_back(B);
    }
    Table tab(c);
    size_t ob = tab.add(boxes);
    if ((rc = tab.upload())) return rc;
    const dim3 gr = grid_for(maxpts, (int)boxes.size(), 1, NTBS);
    if (sph)
        MX_LAUNCH(c, "k_vel_force", k_vel_force<true>, gr, dim3(NTBS), tab.dev<VfBox>(ob), G, A);
    else
        MX_LAUNCH(c, "k_vel_force", k_vel_force<false>, gr, dim3(NTBS), tab.dev<VfBox>(ob), G, A);
    return MX_OK;
}

int mx_adv_modify_scal_force(mx_adv_ctx* c, int lev, mx_multifab* scal_force,
                             const mx_multifab* state, const mx_face_multifab* umac,
                             const mx_basestate* s0_edge, const mx_multifab* s0_cart,
                             const mx_multifab* w0_cart, const mx_basestate* w0, int comp,
                             int fullform, const mx_sph_geom* sg) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    const mx_adv_params& p = c->p;
    const int dim = p.dim;
    const bool sph = dim == 3 && p.spherical;
    if (comp < 0 || comp >= p.nscal) return MX_ERR_ARG;
    if ((rc = check_mf(c, lev, scal_force, comp + 1, 0, -1))) return rc;
    if ((rc = check_mf(c, lev, state, comp + 1, 0, -1))) return rc;
    if ((rc = check_face(c, lev, umac, 1, 0))) return rc;
    SfArgs A;
    for (int d = 0; d < 3; ++d) {
        A.dx[d] = p.dx[lev][d];
        A.domlo[d] = p.domain_lo[lev][d];
        A.domhi[d] = p.domain_hi[lev][d];
    }
    A.comp = comp;
    A.fullform = fullform != 0;
    A.dim = dim;
    A.s0_type = 0;
    A.vlo = p.domain_lo[lev][dim - 1];
    A.vhi = p.domain_hi[lev][dim - 1];
    A.vper = p.is_periodic[dim - 1] != 0;
    A.s0e = nullptr;
    A.divw0 = nullptr;
    if (sph) {
        if (!w0 || !w0->p || !sg) return MX_ERR_ARG;
        if ((rc = sph_check(c, lev, sg, w0, 2, 0))) return rc;
        if (!A.fullform && (rc = check_mf(c, lev, s0_cart, 1, 1, -1))) return rc;
        A.s0_type = sg->s0_interp_type;
        if (A.s0_type < 1 || A.s0_type > 3) return MX_ERR_ARG;
        double* dw = nullptr;
        if ((rc = radial_w0(c, w0, sg, 1, p.use_exact_base_state, &dw))) return rc;
        A.divw0 = dw;
    } else {
        if ((rc = check_mf(c, lev, w0_cart, dim, 1, -1))) return rc;
        if (!A.fullform) {
            if ((rc = check_mf(c, lev, s0_cart, 1, 0, -1))) return rc;
            if (!s0_edge || !s0_edge->p || lev >= s0_edge->nlev || s0_edge->len < A.vhi + 1)
                return MX_ERR_ARG;
            A.s0e = s0_edge->p + (long long)lev * s0_edge->len;
        }
    }
    const SphG G = sph_g(c, lev, sg);
    std::vector<SfBox> boxes;
    long long maxpts = 0;
    for (size_t b = 0; b < c->grids[lev].size(); ++b) {
        SfBox B;
        HBox r = hbox(c->grids[lev][b]);
        for (int d = 0; d < 3; ++d) {
            B.lo[d] = r.lo[d];
            B.hi[d] = r.hi[d];
        }
        B.f = to_da(scal_force->fab[b]);
        B.s = to_da(state->fab[b]);
        B.s0c = A.fullform ? null_da() : to_da(s0_cart->fab[b]);
        B.w0c = sph ? null_da() : to_da(w0_cart->fab[b]);
        for (int d = 0; d < 3; ++d) B.um[d] = d < dim ? to_da(umac->dir[d].fab[b]) : null_da();
        maxpts = std::max(maxpts, points(r));
        boxes.push_back(B);
    }
    Table tab(c);
    size_t ob = tab.add(boxes);
    if ((rc = tab.upload())) return rc;
    const dim3 gr = grid_for(maxpts, (int)boxes.size(), 1, NTBS);
    if (sph)
        MX_LAUNCH(c, "k_scal_force", k_scal_force<true>, gr, dim3(NTBS), tab.dev<SfBox>(ob), G, A);
    else
        MX_LAUNCH(c, "k_scal_force", k_scal_force<false>, gr, dim3(NTBS), tab.dev<SfBox>(ob), G, A);
    return MX_OK;
}

int mx_adv_convert_rhox_to_x(mx_adv_ctx* c, int lev, mx_multifab* scal, int flag) {
    return convert(c, lev, scal, c ? c->p.first_spec : 0, c ? c->p.num_spec : 0, flag);
}

int mx_adv_convert_rhoh_to_h(mx_adv_ctx* c, int lev, mx_multifab* scal, int flag) {
    return convert(c, lev, scal, c ? c->p.comp_rhoh : 0, 1, flag);
}

}  // extern "C"
