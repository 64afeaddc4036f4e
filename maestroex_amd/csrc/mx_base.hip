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

int mx_adv_convert_rhox_to_x(mx_adv_ctx* c, int lev, mx_multifab* scal, int flag) {
    return convert(c, lev, scal, c ? c->p.first_spec : 0, c ? c->p.num_spec : 0, flag);
}

int mx_adv_convert_rhoh_to_h(mx_adv_ctx* c, int lev, mx_multifab* scal, int flag) {
    return convert(c, lev, scal, c ? c->p.comp_rhoh : 0, 1, flag);
}

}  // extern "C"
