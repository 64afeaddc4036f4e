// oracle_api.cpp -- C entry points of the CPU oracle (liboracle.so).
//
// TEST INFRASTRUCTURE ONLY (see oracle_core.H).  The functions mirror the
// mx_adv_* entry points of include/mx_adv.h one for one (prefix mxo_ instead
// of mx_adv_), with every field in host memory, so that tests can run the
// same call sequence on the oracle and on the HIP library and compare.
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_core.H"

namespace mxo {

std::atomic<int> g_oob{0};
std::atomic<long long> g_alloc_ns{0};

void PPM(const Ctx& c, const IBox& bx, const Arr& s, const Arr& u,
         const Arr& v, const Arr& w, const Arr& Ip, const Arr& Im,
         const IBox& dom, const mx_bcrec* bcs, const double* dx, bool is_umac,
         int comp, int bccomp);
void Slope(int dir, int slope_order, const IBox& bx, const Arr& s,
           const Arr& sl, const IBox& dom, const mx_bcrec* bcs, int ncomp,
           int bc_start_comp);
void BCSetup(const Ctx& c, std::vector<mx_bcrec>& bcs_s,
             std::vector<mx_bcrec>& bcs_u);
void FillFaceGhosts(const Ctx& c, int lev, const std::vector<Arr> uedge[3],
                    bool extrapolate = true);
int MakeUtrans(const Ctx& c, int lev, const std::vector<Arr>& utilde,
               const std::vector<Arr>& ufull, const std::vector<Arr> utrans[3],
               const std::vector<Arr> w0mac[3], const std::vector<Arr>& w0_cart,
               const std::vector<mx_bcrec>& bcs_u, bool fill_ghosts = true);
int Addw0(const Ctx& c, int lev, const std::vector<Arr> uedge[3],
          const std::vector<Arr> w0mac[3], const BaseArr& w0, double mult,
          bool fill_ghosts = true);
int FillPatchUedgeCF(const Ctx& c, int lev, const std::vector<Arr> crse[3],
                     const std::vector<Arr> fine[3]);
int VelPred(const Ctx& c, int lev, const std::vector<Arr>& utilde,
            const std::vector<Arr>& ufull, const std::vector<Arr> utrans[3],
            const std::vector<Arr> umac[3], const std::vector<Arr> w0mac[3],
            const std::vector<Arr>& force, const std::vector<Arr>& w0_cart,
            const std::vector<mx_bcrec>& bcs_u);
int MakeEdgeScal(const Ctx& c, int lev, const std::vector<Arr>& state,
                 const std::vector<Arr> sedge[3], const std::vector<Arr> umac[3],
                 const std::vector<Arr>& force, bool is_vel,
                 const mx_bcrec* bcs, int start_scomp, int start_bccomp,
                 int num_comp, bool is_conservative);
int MakeRhoXFlux(const Ctx& c, int lev, const std::vector<Arr> sflux[3],
                 const std::vector<Arr>& etarhoflux, const std::vector<Arr> sedge[3],
                 const std::vector<Arr> umac[3], const BaseArr& rho0_old,
                 const BaseArr& rho0_edge_old, const std::vector<Arr> r0mac_old[3],
                 const BaseArr& rho0_new, const BaseArr& rho0_edge_new,
                 const std::vector<Arr> r0mac_new[3], const BaseArr& rho0_predicted_edge,
                 const BaseArr& w0, int start_comp, int num_comp);
int MakeRhoHFlux(const Ctx& c, int lev, const std::vector<Arr> sflux[3],
                 const std::vector<Arr> sedge[3], const std::vector<Arr> umac[3],
                 const BaseArr& rho0_old, const BaseArr& rho0_edge_old,
                 const std::vector<Arr> r0mac_old[3], const BaseArr& rho0_new,
                 const BaseArr& rho0_edge_new, const std::vector<Arr> r0mac_new[3],
                 const BaseArr& rhoh0_old, const BaseArr& rhoh0_edge_old,
                 const std::vector<Arr> rh0mac_old[3], const BaseArr& rhoh0_new,
                 const BaseArr& rhoh0_edge_new, const std::vector<Arr> rh0mac_new[3],
                 const std::vector<Arr> h0mac_old[3], const std::vector<Arr> h0mac_new[3]);
void CelltoEdge(int lev, int nr, const mx_basestate* cell, mx_basestate* edge);
int MakeEdgeState1d(const Ctx* c, const mx_radial_geom* rg, const mx_basestate* s,
                    mx_basestate* sedge, const mx_basestate* force, const mx_basestate* w0);
void ScalStep(int n, int ncomp, int ppm, double dt, const double dx[3], double rel_eps,
              const double* s, const double* const u[3], double* snew, const double* ds,
              const double* const du[3], double* dsnew);
void ScalStepBC(const int n[3], const int per[3], int ncomp, int ppm, double dt,
                const double dx[3], double rel_eps, const double* s, const double* const u[3],
                double* snew, const double* ds, const double* const du[3], double* dsnew,
                bool cons);
void StepOpCount(int n, double dt, const double dx[3], double rel_eps, const double* s,
                 const double* const u[3], double out[7]);
int Put1dArrayOnCart(const Ctx& c, int lev, const BaseArr& s0, const std::vector<Arr>& s0_cart,
                     bool edge, bool vec, const mx_sph_geom* sg);
int MakeW0mac(const Ctx& c, int lev, const BaseArr& w0, const std::vector<Arr>& w0_cart,
              const std::vector<Arr> w0mac[3], const mx_sph_geom* sg);
int MakeS0mac(const Ctx& c, int lev, const BaseArr& s0, const std::vector<Arr>& s0_cart,
              const std::vector<Arr> s0mac[3], const mx_sph_geom* sg);
int PutInPertForm(const Ctx& c, int lev, const std::vector<Arr>& scal, const BaseArr& s0,
                  int comp, bool flag, const mx_sph_geom* sg);
int ConvertByRho(const Ctx& c, int lev, const std::vector<Arr>& scal, int comp0, int ncomp,
                 bool flag);
int AverageDown(const Ctx& c, int lev, const std::vector<Arr>& fine,
                const std::vector<Arr>& crse, int comp, int ncomp);
int FillPatchTwoLevels(const Ctx& c, int lev, const std::vector<Arr>& mf,
                       const std::vector<Arr>& crse, const std::vector<Arr>& fine, int scomp,
                       int dcomp, int ncomp, const mx_bcrec* bcs, bool is_vel);
int FillPatchSingleLevel(const Ctx& c, int lev, const std::vector<Arr>& mf,
                         const std::vector<Arr>& src, int scomp, int dcomp, int ncomp,
                         const mx_bcrec* bcs, bool is_vel);
int AverageDownFaces(const Ctx& c, int lev, const std::vector<Arr> fine[3],
                     const std::vector<Arr> crse[3]);
int FluxRegDefine(const Ctx& c, int lev, int ncomp);
int FluxRegSetVal(const Ctx& c, int lev, double v);
int FluxRegCrseInit(const Ctx& c, int lev, const std::vector<Arr>& flux, int dir, int scomp,
                    int dcomp, int ncomp, double mult);
int FluxRegFineAdd(const Ctx& c, int lev, const std::vector<Arr>& flux, int dir, int scomp,
                   int dcomp, int ncomp, double mult);
int FluxRegReflux(const Ctx& c, int lev, const std::vector<Arr>& state, double scale, int scomp,
                  int dcomp, int ncomp);
int MakeVelForce(const Ctx& c, int lev, const std::vector<Arr>& vel_force,
                 const std::vector<Arr> uedge[3], const std::vector<Arr>& rho,
                 const std::vector<Arr>& gpi, const std::vector<Arr>& w0_force_cart,
                 const std::vector<Arr>& w0_cart, const std::vector<Arr>& normal,
                 const BaseArr& rho0, const BaseArr& grav_cell, const BaseArr& w0,
                 double buoyancy_cutoff_factor, double base_cutoff_density,
                 bool do_add_utilde_force, bool average_base_state, const mx_sph_geom* sg);
int ModifyScalForce(const Ctx& c, int lev, const std::vector<Arr>& scal_force,
                    const std::vector<Arr>& state, const std::vector<Arr> umac[3],
                    const BaseArr& s0_edge, const std::vector<Arr>& s0_cart,
                    const std::vector<Arr>& w0_cart, const BaseArr& w0, int comp, bool fullform,
                    const mx_sph_geom* sg);
int UpdateScal(const Ctx& c, int lev, const std::vector<Arr>& sold, const std::vector<Arr>& snew,
               const std::vector<Arr> sflux[3], const std::vector<Arr>& force, int start_comp,
               double base_cutoff_density);
int UpdateVel(const Ctx& c, int lev, const std::vector<Arr>& uold, const std::vector<Arr>& unew,
              const std::vector<Arr> umac[3], const std::vector<Arr> uedge[3],
              const std::vector<Arr>& force, const std::vector<Arr>& sponge,
              const std::vector<Arr> w0mac[3], const std::vector<Arr>& w0_cart, bool do_sponge);

}  // namespace mxo

using namespace mxo;

namespace {

struct Face {
    std::vector<Arr> d[3];
};
Face faces(const mx_face_multifab* f) {
    Face r;
    if (f)
        for (int d = 0; d < 3; ++d) r.d[d] = arrs(&f->dir[d]);
    return r;
}

int finish(int rc) {
    if (g_oob.exchange(0) != 0) return MX_ERR_BOX;
    return rc;
}

int check_level(const Ctx* c, int lev) {
    if (!c || lev < 0 || lev > c->p.finest_level) return MX_ERR_ARG;
    if (c->grids[lev].empty()) return MX_ERR_ARG;
    return MX_OK;
}

}  // namespace

extern "C" {

int mxo_create(const mx_adv_params* p, Ctx** out) {
    if (!p || !out) return MX_ERR_ARG;
    if (p->dim != 2 && p->dim != 3) return MX_ERR_ARG;
    if (p->ppm_type < 0 || p->ppm_type > 2) return MX_ERR_ARG;
    if (p->finest_level < 0 || p->finest_level >= MX_MAX_LEV) return MX_ERR_ARG;
    if (p->nscal <= 0 || p->nscal > MX_MAX_SCAL) return MX_ERR_ARG;
    if (p->spherical && p->dim != 3) return MX_ERR_ARG;
    Ctx* c = new Ctx();
    c->p = *p;
    *out = c;
    return MX_OK;
}

int mxo_destroy(Ctx* c) {
    delete c;
    return MX_OK;
}

int mxo_set_level(Ctx* c, int lev, int nbox, const mx_box* boxes) {
    if (!c || lev < 0 || lev > c->p.finest_level || nbox <= 0 || !boxes) return MX_ERR_ARG;
    c->grids[lev].clear();
    for (int b = 0; b < nbox; ++b) {
        IBox bx = make_box(boxes[b].lo, boxes[b].hi);
        if (c->p.dim == 2) bx.lo[2] = bx.hi[2] = 0;
        c->grids[lev].push_back(bx);
    }
    // a BoxArray's boxes are disjoint
    for (int a = 0; a < nbox; ++a)
        for (int b = a + 1; b < nbox; ++b) {
            const IBox &x = c->grids[lev][a], &y = c->grids[lev][b];
            bool overlap = true;
            for (int d = 0; d < 3; ++d)
                if ((x.lo[d] > y.lo[d] ? x.lo[d] : y.lo[d]) > (x.hi[d] < y.hi[d] ? x.hi[d] : y.hi[d]))
                    overlap = false;
            if (overlap) {
                c->grids[lev].clear();
                return MX_ERR_BOX;
            }
        }
    // reflux_type 2: the level's flux register, Nscal components
    // (MaestroInit.cpp:389-392)
    if (lev > 0 && c->p.reflux_type == 2) return FluxRegDefine(*c, lev, c->p.nscal);
    return MX_OK;
}

int mxo_set_dt(Ctx* c, double dt, double rel_eps) {
    if (!c) return MX_ERR_ARG;
    c->p.dt = dt;
    c->p.rel_eps = rel_eps;
    return MX_OK;
}

int mxo_bc_setup(const Ctx* c, mx_bcrec* bcs_s, mx_bcrec* bcs_u) {
    if (!c) return MX_ERR_ARG;
    std::vector<mx_bcrec> s, u;
    BCSetup(*c, s, u);
    if (bcs_s) std::memcpy(bcs_s, s.data(), sizeof(mx_bcrec) * s.size());
    if (bcs_u) std::memcpy(bcs_u, u.data(), sizeof(mx_bcrec) * c->p.dim);
    return MX_OK;
}

int mxo_ppm(Ctx* c, int lev, const mx_box* bx, const mx_fab* s, const mx_fab* u,
            const mx_fab* v, const mx_fab* w, mx_fab* Ip, mx_fab* Im,
            const mx_bcrec* bcs, int nbcs, int is_umac, int comp, int bccomp) {
    if (!c || !bx || !s || !u || !v || !Ip || !Im || !bcs) return MX_ERR_ARG;
    if (c->p.dim == 3 && !w) return MX_ERR_ARG;
    if (bccomp < 0 || bccomp >= nbcs) return MX_ERR_ARG;
    IBox b = make_box(bx->lo, bx->hi);
    if (c->p.dim == 2) b.lo[2] = b.hi[2] = 0;
    Arr wa;
    if (w) wa = from_fab(*w);
    PPM(*c, b, from_fab(*s), from_fab(*u), from_fab(*v), wa, from_fab(*Ip),
        from_fab(*Im), c->domain(lev), bcs, c->p.dx[lev], is_umac != 0, comp,
        bccomp);
    return finish(MX_OK);
}

int mxo_slope(Ctx* c, int lev, int dir, const mx_box* bx, const mx_fab* s,
              mx_fab* sl, const mx_bcrec* bcs, int nbcs, int ncomp,
              int bc_start_comp, int slope_order) {
    if (!c || !bx || !s || !sl || !bcs || dir < 0 || dir >= c->p.dim) return MX_ERR_ARG;
    if (bc_start_comp < 0 || bc_start_comp + ncomp > nbcs) return MX_ERR_ARG;
    IBox b = make_box(bx->lo, bx->hi);
    if (c->p.dim == 2) b.lo[2] = b.hi[2] = 0;
    Slope(dir, slope_order, b, from_fab(*s), from_fab(*sl), c->domain(lev), bcs,
          ncomp, bc_start_comp);
    return finish(MX_OK);
}

int mxo_make_utrans(Ctx* c, int lev, const mx_multifab* utilde,
                    const mx_multifab* ufull, mx_face_multifab* utrans,
                    const mx_face_multifab* w0mac, const mx_multifab* w0_cart) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    std::vector<mx_bcrec> bs, bu;
    BCSetup(*c, bs, bu);
    Face tr = faces(utrans), wm = faces(w0mac);
    rc = MakeUtrans(*c, lev, arrs(utilde), arrs(ufull), tr.d, wm.d, arrs(w0_cart), bu);
    return finish(rc);
}

int mxo_addw0(Ctx* c, int lev, mx_face_multifab* uedge, const mx_face_multifab* w0mac,
              const mx_basestate* w0, double mult) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    Face ue = faces(uedge), wm = faces(w0mac);
    rc = Addw0(*c, lev, ue.d, wm.d, base(w0), mult);
    return finish(rc);
}

int mxo_fill_face_ghosts(Ctx* c, int lev, mx_face_multifab* uedge) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    Face ue = faces(uedge);
    FillFaceGhosts(*c, lev, ue.d);
    return finish(MX_OK);
}

int mxo_velpred(Ctx* c, int lev, const mx_multifab* utilde, const mx_multifab* ufull,
                const mx_face_multifab* utrans, mx_face_multifab* umac,
                const mx_face_multifab* w0mac, const mx_multifab* force,
                const mx_multifab* w0_cart) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    std::vector<mx_bcrec> bs, bu;
    BCSetup(*c, bs, bu);
    Face tr = faces(utrans), um = faces(umac), wm = faces(w0mac);
    rc = VelPred(*c, lev, arrs(utilde), arrs(ufull), tr.d, um.d, wm.d, arrs(force),
                 arrs(w0_cart), bu);
    return finish(rc);
}

int mxo_make_edge_scal(Ctx* c, int lev, const mx_multifab* state, mx_face_multifab* sedge,
                       const mx_face_multifab* umac, const mx_multifab* force, int is_vel,
                       const mx_bcrec* bcs, int nbccomp, int start_scomp, int start_bccomp,
                       int num_comp, int is_conservative) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    if (!bcs || start_bccomp < 0 || start_bccomp + num_comp > nbccomp) return MX_ERR_ARG;
    Face se = faces(sedge), um = faces(umac);
    rc = MakeEdgeScal(*c, lev, arrs(state), se.d, um.d, arrs(force), is_vel != 0, bcs,
                      start_scomp, start_bccomp, num_comp, is_conservative != 0);
    return finish(rc);
}

int mxo_make_rhox_flux(Ctx* c, int lev, mx_face_multifab* sflux, mx_multifab* etarhoflux,
                       const mx_face_multifab* sedge, const mx_face_multifab* umac,
                       const mx_basestate* rho0_old, const mx_basestate* rho0_edge_old,
                       const mx_face_multifab* r0mac_old, const mx_basestate* rho0_new,
                       const mx_basestate* rho0_edge_new, const mx_face_multifab* r0mac_new,
                       const mx_basestate* rho0_predicted_edge, const mx_basestate* w0,
                       int start_comp, int num_comp) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    Face sf = faces(sflux), se = faces(sedge), um = faces(umac), ro = faces(r0mac_old),
         rn = faces(r0mac_new);
    rc = MakeRhoXFlux(*c, lev, sf.d, arrs(etarhoflux), se.d, um.d, base(rho0_old),
                      base(rho0_edge_old), ro.d, base(rho0_new), base(rho0_edge_new), rn.d,
                      base(rho0_predicted_edge), base(w0), start_comp, num_comp);
    return finish(rc);
}

int mxo_make_rhoh_flux(Ctx* c, int lev, mx_face_multifab* sflux, const mx_face_multifab* sedge,
                       const mx_face_multifab* umac, const mx_face_multifab* w0mac,
                       const mx_basestate* rho0_old, const mx_basestate* rho0_edge_old,
                       const mx_face_multifab* r0mac_old, const mx_basestate* rho0_new,
                       const mx_basestate* rho0_edge_new, const mx_face_multifab* r0mac_new,
                       const mx_basestate* rhoh0_old, const mx_basestate* rhoh0_edge_old,
                       const mx_face_multifab* rh0mac_old, const mx_basestate* rhoh0_new,
                       const mx_basestate* rhoh0_edge_new, const mx_face_multifab* rh0mac_new,
                       const mx_face_multifab* h0mac_old, const mx_face_multifab* h0mac_new) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    (void)w0mac;
    Face sf = faces(sflux), se = faces(sedge), um = faces(umac), ro = faces(r0mac_old),
         rn = faces(r0mac_new), rho_ = faces(rh0mac_old), rhn = faces(rh0mac_new),
         ho = faces(h0mac_old), hn = faces(h0mac_new);
    rc = MakeRhoHFlux(*c, lev, sf.d, se.d, um.d, base(rho0_old), base(rho0_edge_old), ro.d,
                      base(rho0_new), base(rho0_edge_new), rn.d, base(rhoh0_old),
                      base(rhoh0_edge_old), rho_.d, base(rhoh0_new), base(rhoh0_edge_new),
                      rhn.d, ho.d, hn.d);
    return finish(rc);
}

int mxo_cell_to_edge(Ctx* c, int lev, int nr, const mx_basestate* s0_cell,
                     mx_basestate* s0_edge) {
    if (!c || !s0_cell || !s0_edge || nr < 4) return MX_ERR_ARG;
    CelltoEdge(lev, nr, s0_cell, s0_edge);
    return MX_OK;
}

// argument checks of mx_adv_make_edge_state_1d (include/mx_adv.h)
static int edge1d_check(const Ctx* c, const mx_radial_geom* rg, const mx_basestate* s,
                        const mx_basestate* sedge, const mx_basestate* force,
                        const mx_basestate* w0) {
    if (!c || !rg || !s || !sedge || !force || !w0) return MX_ERR_ARG;
    if (!s->p || !sedge->p || !force->p || !w0->p) return MX_ERR_ARG;
    const int nlev = c->p.spherical ? 1 : rg->nlev;
    if (nlev < 1 || nlev > MX_MAX_LEV || rg->nr_fine < 4) return MX_ERR_ARG;
    if (s->len < rg->nr_fine || force->len < rg->nr_fine || sedge->len < rg->nr_fine + 1 ||
        w0->len < rg->nr_fine + 1)
        return MX_ERR_ARG;
    if (s->nlev < nlev || force->nlev < nlev || sedge->nlev < nlev || w0->nlev < nlev)
        return MX_ERR_ARG;
    if (c->p.spherical) return rg->dr[0] > 0.0 ? MX_OK : MX_ERR_ARG;
    for (int n = 0; n < nlev; ++n) {
        const int nr = rg->nr[n];
        if (nr < 4 || nr > rg->nr_fine || !(rg->dr[n] > 0.0)) return MX_ERR_ARG;
        if (rg->nchunks[n] < 0 || rg->nchunks[n] > MX_MAX_CHUNK) return MX_ERR_ARG;
        if (n + 1 < nlev && rg->nchunks[n + 1] < rg->nchunks[n]) return MX_ERR_UNSUPPORTED;
        for (int i = 0; i < rg->nchunks[n]; ++i) {
            const int lo = rg->r_start[n][i], hi = rg->r_end[n][i];
            if (lo < 0 || hi < lo || hi > nr - 1) return MX_ERR_ARG;
            // "make_edge_state assumes blocking_factor > 1" (MaestroMakeEdgeState.cpp:427-436)
            if ((c->p.ppm_type == 1 || c->p.ppm_type == 2) && (lo == 2 || hi == nr - 3))
                return MX_ERR_ARG;
            if (n > 0 && (lo / 2 > rg->nr[n - 1] || (hi + 1) / 2 > rg->nr[n - 1]))
                return MX_ERR_ARG;
        }
    }
    return MX_OK;
}

int mxo_make_edge_state_1d(Ctx* c, const mx_radial_geom* rg, const mx_basestate* s,
                           mx_basestate* sedge, const mx_basestate* force,
                           const mx_basestate* w0) {
    int rc = edge1d_check(c, rg, s, sedge, force, w0);
    if (rc) return rc;
    return MakeEdgeState1d(c, rg, s, sedge, force, w0);
}

// The differentiable scalar step on a periodic box (dense N^3 arrays, face N
// is face 0) and its Jacobian-vector product (du[d] may be null).
int mxo_scal_step(int n, int ncomp, int ppm, double dt, const double* dx, double rel_eps,
                  const double* s, const double* u0, const double* u1, const double* u2,
                  double* snew) {
    if (n < 4 || ncomp < 1 || (ppm != 1 && ppm != 2) || !dx || !s || !u0 || !u1 || !u2 || !snew)
        return MX_ERR_ARG;
    const double* u[3] = {u0, u1, u2};
    const double* du[3] = {nullptr, nullptr, nullptr};
    ScalStep(n, ncomp, ppm, dt, dx, rel_eps, s, u, snew, nullptr, du, nullptr);
    return MX_OK;
}

int mxo_scal_step_jvp(int n, int ncomp, int ppm, double dt, const double* dx, double rel_eps,
                      const double* s, const double* u0, const double* u1, const double* u2,
                      const double* ds, const double* du0, const double* du1,
                      const double* du2, double* snew, double* dsnew) {
    if (n < 4 || ncomp < 1 || (ppm != 1 && ppm != 2) || !dx || !s || !u0 || !u1 || !u2 || !ds ||
        !snew || !dsnew)
        return MX_ERR_ARG;
    const double* u[3] = {u0, u1, u2};
    const double* du[3] = {du0, du1, du2};
    ScalStep(n, ncomp, ppm, dt, dx, rel_eps, s, u, snew, ds, du, dsnew);
    return MX_OK;
}

// The step on n[0] x n[1] x n[2] cells, each direction periodic (per[d])
// or outflow (face arrays one point longer along their outflow direction),
// advective or conservative form, and its Jacobian-vector product when ds
// is given (du[d] may be null).
int mxo_scal_step_bc(const int* n, const int* per, int ncomp, int ppm, double dt,
                     const double* dx, double rel_eps, const double* s, const double* u0,
                     const double* u1, const double* u2, double* snew, const double* ds,
                     const double* du0, const double* du1, const double* du2, double* dsnew,
                     int is_conservative) {
    if (!n || !per || ncomp < 1 || (ppm != 1 && ppm != 2) || !dx || !s || !u0 || !u1 || !u2 ||
        !snew || (ds && !dsnew))
        return MX_ERR_ARG;
    for (int d = 0; d < 3; ++d)
        if (n[d] < 4) return MX_ERR_ARG;
    const double* u[3] = {u0, u1, u2};
    const double* du[3] = {du0, du1, du2};
    ScalStepBC(n, per, ncomp, ppm, dt, dx, rel_eps, s, u, snew, ds, du, dsnew,
               is_conservative != 0);
    return MX_OK;
}

// FP64 operation counts per cell of the step of one component (see
// StepOpCount in oracle_step.cpp).
int mxo_step_op_count(int n, double dt, const double* dx, double rel_eps, const double* s,
                      const double* u0, const double* u1, const double* u2, double* out) {
    if (n < 4 || !dx || !s || !u0 || !u1 || !u2 || !out) return MX_ERR_ARG;
    const double* u[3] = {u0, u1, u2};
    StepOpCount(n, dt, dx, rel_eps, s, u, out);
    return MX_OK;
}

// The pass of mx_adv_advance_pass (include/mx_adv.h).
int mxo_advance_pass(Ctx* c, int lev, const mx_adv_pass_args* a) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    if (!a) return MX_ERR_ARG;
    std::vector<mx_bcrec> bs, bu;
    BCSetup(*c, bs, bu);
    const mx_adv_params& p = c->p;

    Face utrans = faces(&a->utrans), umac = faces(&a->umac), sedge = faces(&a->sedge),
         sflux = faces(&a->sflux), w0mac = faces(&a->w0mac), none;
    const bool sph = c->p.dim == 3 && c->p.spherical;
    Face r0mac, rh0mac, h0mac;
    if (sph) {
        r0mac = faces(&a->r0mac);
        rh0mac = faces(&a->rh0mac);
        h0mac = faces(&a->h0mac);
    }
    auto utilde = arrs(&a->utilde), ufull = arrs(&a->ufull), vel_force = arrs(&a->vel_force),
         w0_cart = arrs(&a->w0_cart), scal = arrs(&a->scal), scal_force = arrs(&a->scal_force),
         eta = arrs(&a->etarhoflux);

    rc = MakeUtrans(*c, lev, utilde, ufull, utrans.d, w0mac.d, w0_cart, bu);
    if (rc) return finish(rc);
    rc = Addw0(*c, lev, utrans.d, w0mac.d, base(&a->w0), 1.0);
    if (rc) return finish(rc);
    rc = VelPred(*c, lev, utilde, ufull, utrans.d, umac.d, w0mac.d, vel_force, w0_cart, bu);
    if (rc) return finish(rc);
    rc = Addw0(*c, lev, umac.d, w0mac.d, base(&a->w0), 1.0);
    if (rc) return finish(rc);
    rc = MakeEdgeScal(*c, lev, scal, sedge.d, umac.d, scal_force, false, bs.data(),
                      p.first_spec, p.first_spec, p.num_spec, false);
    if (rc) return finish(rc);
    rc = MakeEdgeScal(*c, lev, scal, sedge.d, umac.d, scal_force, false, bs.data(),
                      p.comp_rho, p.comp_rho, 1, false);
    if (rc) return finish(rc);
    rc = MakeEdgeScal(*c, lev, scal, sedge.d, umac.d, scal_force, false, bs.data(),
                      p.comp_rhoh, p.comp_rhoh, 1, false);
    if (rc) return finish(rc);
    // etarhoflux[lev].setVal(0.) (MaestroAdvance.cpp:442-445), ghosts included
    for (size_t b = 0; b < eta.size(); ++b) {
        const Arr& e = eta[b];
        IBox whole;
        for (int d = 0; d < 3; ++d) {
            whole.lo[d] = e.b[d];
            whole.hi[d] = e.e[d] - 1;
        }
        par_for(whole, [&](int i, int j, int k) { e(i, j, k) = 0.0; });
    }
    rc = MakeRhoXFlux(*c, lev, sflux.d, eta, sedge.d, umac.d, base(&a->rho0),
                      base(&a->rho0_edge), r0mac.d, base(&a->rho0), base(&a->rho0_edge), r0mac.d,
                      base(&a->rho0_predicted_edge), base(&a->w0), p.first_spec, p.num_spec);
    if (rc) return finish(rc);
    rc = MakeRhoHFlux(*c, lev, sflux.d, sedge.d, umac.d, base(&a->rho0), base(&a->rho0_edge),
                      r0mac.d, base(&a->rho0), base(&a->rho0_edge), r0mac.d, base(&a->rhoh0),
                      base(&a->rhoh0_edge), rh0mac.d, base(&a->rhoh0), base(&a->rhoh0_edge),
                      rh0mac.d, h0mac.d, h0mac.d);
    return finish(rc);
}

// The pass over levels 0..nlev-1 (mx_adv_advance_pass_levels), in the
// reference's order with finest_level = nlev - 1 > 0: every stage loops over
// the levels, and the face data are synchronised between levels where the
// reference does it (reflux_type 1, the default, _cpp_parameters:24):
//   MakeUtrans; AverageDownFaces(utrans); FillPatchUedge(utrans)
//     (MaestroMakeUtrans.cpp:641-647)
//   Addw0(utrans, +1); AverageDownFaces; FillPatchUedge (MaestroFill3dData.cpp:449-455)
//   VelPred; AverageDownFaces(umac) (MaestroVelPred.cpp:291)
//   Addw0(umac, +1); AverageDownFaces; FillPatchUedge
//   MakeEdgeScal x 3 (scalars: no restriction, MaestroMakeEdgeScalars.cpp:294-301)
//   MakeRhoXFlux; AverageDownFaces(sflux) (MaestroMakeFlux.cpp:459-462)
//   MakeRhoHFlux; AverageDownFaces(sflux) (MaestroMakeFlux.cpp:1061-1064)
// nlev == 1 is the single-level pass.
int mxo_advance_pass_levels(Ctx* c, int nlev, const mx_adv_pass_args* a) {
    if (!c || !a || nlev < 1 || nlev - 1 > c->p.finest_level) return MX_ERR_ARG;
    if (nlev == 1) return mxo_advance_pass(c, 0, a);
    for (int lev = 0; lev < nlev; ++lev) {
        int rc = check_level(c, lev);
        if (rc) return rc;
    }
    std::vector<mx_bcrec> bs, bu;
    BCSetup(*c, bs, bu);
    const mx_adv_params& p = c->p;
    const bool sph = p.dim == 3 && p.spherical;
    std::vector<Face> utrans(nlev), umac(nlev), sedge(nlev), sflux(nlev), w0mac(nlev), r0mac(nlev),
        rh0mac(nlev), h0mac(nlev);
    for (int l = 0; l < nlev; ++l) {
        utrans[l] = faces(&a[l].utrans);
        umac[l] = faces(&a[l].umac);
        sedge[l] = faces(&a[l].sedge);
        sflux[l] = faces(&a[l].sflux);
        w0mac[l] = faces(&a[l].w0mac);
        if (sph) {
            r0mac[l] = faces(&a[l].r0mac);
            rh0mac[l] = faces(&a[l].rh0mac);
            h0mac[l] = faces(&a[l].h0mac);
        }
    }
    auto average_down_faces = [&](std::vector<Face>& f) {
        for (int lev = nlev - 2; lev >= 0; --lev) AverageDownFaces(*c, lev, f[lev + 1].d, f[lev].d);
    };
    // Maestro::FillPatchUedge (MaestroFillData.cpp:455-833)
    auto fill_patch_uedge = [&](std::vector<Face>& f) {
        for (int lev = 0; lev < nlev; ++lev) {
            if (lev > 0) FillPatchUedgeCF(*c, lev, f[lev - 1].d, f[lev].d);
            FillFaceGhosts(*c, lev, f[lev].d, false);
        }
    };
    (void)fill_patch_uedge;
    int rc;
    for (int lev = 0; lev < nlev; ++lev) {
        rc = MakeUtrans(*c, lev, arrs(&a[lev].utilde), arrs(&a[lev].ufull), utrans[lev].d,
                        w0mac[lev].d, arrs(&a[lev].w0_cart), bu, false);
        if (rc) return finish(rc);
    }
    average_down_faces(utrans);
    fill_patch_uedge(utrans);
    for (int lev = 0; lev < nlev; ++lev) {
        rc = Addw0(*c, lev, utrans[lev].d, w0mac[lev].d, base(&a[lev].w0), 1.0, false);
        if (rc) return finish(rc);
    }
    average_down_faces(utrans);
    fill_patch_uedge(utrans);
    for (int lev = 0; lev < nlev; ++lev) {
        rc = VelPred(*c, lev, arrs(&a[lev].utilde), arrs(&a[lev].ufull), utrans[lev].d,
                     umac[lev].d, w0mac[lev].d, arrs(&a[lev].vel_force), arrs(&a[lev].w0_cart), bu);
        if (rc) return finish(rc);
    }
    average_down_faces(umac);
    for (int lev = 0; lev < nlev; ++lev) {
        rc = Addw0(*c, lev, umac[lev].d, w0mac[lev].d, base(&a[lev].w0), 1.0, false);
        if (rc) return finish(rc);
    }
    average_down_faces(umac);
    fill_patch_uedge(umac);
    const int groups[3][2] = {{p.first_spec, p.num_spec}, {p.comp_rho, 1}, {p.comp_rhoh, 1}};
    for (const auto& gr : groups)
        for (int lev = 0; lev < nlev; ++lev) {
            rc = MakeEdgeScal(*c, lev, arrs(&a[lev].scal), sedge[lev].d, umac[lev].d,
                              arrs(&a[lev].scal_force), false, bs.data(), gr[0], gr[0], gr[1],
                              false);
            if (rc) return finish(rc);
        }
    for (int lev = 0; lev < nlev; ++lev) {
        auto eta = arrs(&a[lev].etarhoflux);
        for (size_t b = 0; b < eta.size(); ++b) {
            const Arr& e = eta[b];
            IBox whole;
            for (int d = 0; d < 3; ++d) {
                whole.lo[d] = e.b[d];
                whole.hi[d] = e.e[d] - 1;
            }
            par_for(whole, [&](int i, int j, int k) { e(i, j, k) = 0.0; });
        }
    }
    for (int lev = 0; lev < nlev; ++lev) {
        const mx_adv_pass_args* A = &a[lev];
        rc = MakeRhoXFlux(*c, lev, sflux[lev].d, arrs(&A->etarhoflux), sedge[lev].d, umac[lev].d,
                          base(&A->rho0), base(&A->rho0_edge), r0mac[lev].d, base(&A->rho0),
                          base(&A->rho0_edge), r0mac[lev].d, base(&A->rho0_predicted_edge),
                          base(&A->w0), p.first_spec, p.num_spec);
        if (rc) return finish(rc);
    }
    if (p.reflux_type == 1) average_down_faces(sflux);
    for (int lev = 0; lev < nlev; ++lev) {
        const mx_adv_pass_args* A = &a[lev];
        rc = MakeRhoHFlux(*c, lev, sflux[lev].d, sedge[lev].d, umac[lev].d, base(&A->rho0),
                          base(&A->rho0_edge), r0mac[lev].d, base(&A->rho0), base(&A->rho0_edge),
                          r0mac[lev].d, base(&A->rhoh0), base(&A->rhoh0_edge), rh0mac[lev].d,
                          base(&A->rhoh0), base(&A->rhoh0_edge), rh0mac[lev].d, h0mac[lev].d,
                          h0mac[lev].d);
        if (rc) return finish(rc);
    }
    if (p.reflux_type == 1) average_down_faces(sflux);
    return finish(MX_OK);
}

// Maestro::FillPatch on level lev > 0 (FillPatchTwoLevels with
// cell_cons_interp): crse is level lev - 1's source, fine level lev's
int mxo_fill_patch_two_levels(Ctx* c, int lev, mx_multifab* mf, const mx_multifab* crse,
                              const mx_multifab* fine, int srccomp, int destcomp, int ncomp,
                              const mx_bcrec* bcs, int variable_type) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    if (lev < 1 || c->grids[lev - 1].empty()) return MX_ERR_ARG;
    const size_t nb = c->grids[lev].size(), ncb = c->grids[lev - 1].size();
    if (!mf || !crse || !fine || !bcs || mf->nfab != (int)nb || fine->nfab != (int)nb ||
        crse->nfab != (int)ncb || srccomp < 0 || destcomp < 0 || ncomp < 1 ||
        ncomp > MX_MAX_SCAL)
        return MX_ERR_ARG;
    std::vector<Arr> m = arrs(mf), f = arrs(fine), cr = arrs(crse);
    if (m[0].p == f[0].p && srccomp != destcomp && srccomp < destcomp + ncomp &&
        destcomp < srccomp + ncomp)
        return MX_ERR_ARG;
    for (size_t b = 0; b < ncb; ++b) {
        const IBox& g = c->grids[lev - 1][b];
        if (!cr[b].valid() || cr[b].nc < srccomp + ncomp) return MX_ERR_ARG;
        for (int d = 0; d < 3; ++d)
            if (cr[b].b[d] > g.lo[d] || cr[b].e[d] - 1 < g.hi[d]) return MX_ERR_BOX;
    }
    for (size_t b = 0; b < nb; ++b) {
        const IBox& g = c->grids[lev][b];
        if (!m[b].valid() || !f[b].valid() || m[b].nc < destcomp + ncomp ||
            f[b].nc < srccomp + ncomp)
            return MX_ERR_ARG;
        for (int d = 0; d < 3; ++d) {
            const int ng = g.lo[d] - m[b].b[d];
            if (ng < 0 || m[b].e[d] - 1 - g.hi[d] != ng || f[b].b[d] > g.lo[d] ||
                f[b].e[d] - 1 < g.hi[d])
                return MX_ERR_BOX;
            if (b > 0 && ng != c->grids[lev][0].lo[d] - m[0].b[d]) return MX_ERR_BOX;
        }
    }
    return finish(FillPatchTwoLevels(*c, lev, m, cr, f, srccomp, destcomp, ncomp, bcs,
                                     variable_type == 1));
}

// amrex::FluxRegister of level lev (Maestro::flux_reg_s[lev])
int mxo_flux_reg_define(Ctx* c, int lev, int ncomp) {
    if (!c) return MX_ERR_ARG;
    return FluxRegDefine(*c, lev, ncomp);
}

static int reg_ok(const Ctx* c, int lev) {
    if (!c || lev < 1 || lev > c->p.finest_level || lev >= MX_MAX_LEV) return MX_ERR_ARG;
    return c->freg[lev].ncomp > 0 ? MX_OK : MX_ERR_ARG;
}

int mxo_flux_reg_setval(Ctx* c, int lev, double v) {
    int rc = reg_ok(c, lev);
    return rc ? rc : FluxRegSetVal(*c, lev, v);
}

static int dir_flux_ok(const Ctx* c, int lev, const std::vector<Arr>& f, int dir, int nc) {
    if (f.size() != c->grids[lev].size()) return MX_ERR_ARG;
    for (size_t b = 0; b < f.size(); ++b) {
        const IBox r = nodal(c->grids[lev][b], dir);
        if (!f[b].valid() || f[b].nc < nc) return MX_ERR_ARG;
        for (int d = 0; d < 3; ++d)
            if (f[b].b[d] > r.lo[d] || f[b].e[d] - 1 < r.hi[d]) return MX_ERR_BOX;
    }
    return MX_OK;
}

int mxo_flux_reg_crse_init(Ctx* c, int lev, const mx_multifab* flux, int dir, int srccomp,
                           int destcomp, int numcomp, double mult) {
    int rc = reg_ok(c, lev);
    if (rc) return rc;
    if (dir < 0 || dir >= c->p.dim || srccomp < 0 || destcomp < 0 || numcomp < 1 ||
        destcomp + numcomp > c->freg[lev].ncomp)
        return MX_ERR_ARG;
    std::vector<Arr> f = arrs(flux);
    if ((rc = dir_flux_ok(c, lev - 1, f, dir, srccomp + numcomp))) return rc;
    return finish(FluxRegCrseInit(*c, lev, f, dir, srccomp, destcomp, numcomp, mult));
}

int mxo_flux_reg_fine_add(Ctx* c, int lev, const mx_multifab* flux, int dir, int srccomp,
                          int destcomp, int numcomp, double mult) {
    int rc = reg_ok(c, lev);
    if (rc) return rc;
    if (dir < 0 || dir >= c->p.dim || srccomp < 0 || destcomp < 0 || numcomp < 1 ||
        destcomp + numcomp > c->freg[lev].ncomp)
        return MX_ERR_ARG;
    std::vector<Arr> f = arrs(flux);
    if ((rc = dir_flux_ok(c, lev, f, dir, srccomp + numcomp))) return rc;
    return finish(FluxRegFineAdd(*c, lev, f, dir, srccomp, destcomp, numcomp, mult));
}

int mxo_flux_reg_reflux(Ctx* c, int lev, mx_multifab* state, double scale, int scomp, int dcomp,
                        int ncomp) {
    int rc = reg_ok(c, lev);
    if (rc) return rc;
    if (scomp < 0 || dcomp < 0 || ncomp < 1 || scomp + ncomp > c->freg[lev].ncomp)
        return MX_ERR_ARG;
    std::vector<Arr> s = arrs(state);
    if (s.size() != c->grids[lev - 1].size()) return MX_ERR_ARG;
    for (size_t b = 0; b < s.size(); ++b)
        if (!s[b].valid() || s[b].nc < dcomp + ncomp) return MX_ERR_ARG;
    return finish(FluxRegReflux(*c, lev, s, scale, scomp, dcomp, ncomp));
}

int mxo_flux_reg_read(Ctx* c, int lev, double* host, long long cap, long long* n) {
    int rc = reg_ok(c, lev);
    if (rc) return rc;
    if (!n) return MX_ERR_ARG;
    long long total = 0;
    for (const auto& v : c->freg[lev].val) total += (long long)v.size();
    *n = total;
    if (!host) return MX_OK;
    if (cap < total) return MX_ERR_ARG;
    for (const auto& v : c->freg[lev].val) {
        std::copy(v.begin(), v.end(), host);
        host += v.size();
    }
    return MX_OK;
}

// Maestro::FillPatchUedge (MaestroFillData.cpp:455-833) over levels
// 0..nlev-1; uedge[lev] holds level lev's face data (one component)
int mxo_fill_patch_uedge(Ctx* c, int nlev, mx_face_multifab* uedge) {
    if (!c || !uedge || nlev < 1 || nlev - 1 > c->p.finest_level) return MX_ERR_ARG;
    std::vector<Face> f(nlev);
    for (int lev = 0; lev < nlev; ++lev) {
        int rc = check_level(c, lev);
        if (rc) return rc;
        f[lev] = faces(&uedge[lev]);
    }
    for (int lev = 0; lev < nlev; ++lev) {
        if (lev > 0) FillPatchUedgeCF(*c, lev, f[lev - 1].d, f[lev].d);
        FillFaceGhosts(*c, lev, f[lev].d, nlev == 1);
    }
    return finish(MX_OK);
}

int mxo_update_scal(Ctx* c, int lev, const mx_multifab* stateold, mx_multifab* statenew,
                    const mx_face_multifab* sflux, const mx_multifab* force, int start_comp,
                    int num_comp, const mx_multifab* p0_cart, double base_cutoff_density,
                    int do_eos_h_above_cutoff) {
    (void)p0_cart;
    int rc = check_level(c, lev);
    if (rc) return rc;
    if (!stateold || !statenew || !sflux || !force) return MX_ERR_ARG;
    if (do_eos_h_above_cutoff) return MX_ERR_UNSUPPORTED;
    if (start_comp == c->p.comp_rhoh) {
        if (num_comp != 1) return MX_ERR_ARG;
    } else if (start_comp == c->p.first_spec) {
        if (num_comp != c->p.num_spec) return MX_ERR_ARG;
    } else {
        return MX_ERR_ARG;
    }
    Face sf = faces(sflux);
    rc = UpdateScal(*c, lev, arrs(stateold), arrs(statenew), sf.d, arrs(force), start_comp,
                    base_cutoff_density);
    return finish(rc);
}

int mxo_update_vel(Ctx* c, int lev, const mx_multifab* uold, mx_multifab* unew,
                   const mx_face_multifab* umac, const mx_face_multifab* uedge,
                   const mx_multifab* force, const mx_multifab* sponge,
                   const mx_face_multifab* w0mac, const mx_multifab* w0_cart, int do_sponge) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    if (!uold || !unew || !umac || !uedge || !force) return MX_ERR_ARG;
    if (do_sponge && (!sponge || sponge->nfab != (int)c->grids[lev].size())) return MX_ERR_ARG;
    if (c->p.spherical) {
        if (!w0mac) return MX_ERR_ARG;
        for (int d = 0; d < 3; ++d)
            if (w0mac->dir[d].nfab != (int)c->grids[lev].size()) return MX_ERR_ARG;
    } else if (!w0_cart || w0_cart->nfab != (int)c->grids[lev].size()) {
        return MX_ERR_ARG;
    }
    Face um = faces(umac), ue = faces(uedge), wm = faces(w0mac);
    std::vector<Arr> sp;
    if (do_sponge) sp = arrs(sponge);
    rc = UpdateVel(*c, lev, arrs(uold), arrs(unew), um.d, ue.d, arrs(force), sp, wm.d,
                   arrs(w0_cart), do_sponge != 0);
    return finish(rc);
}

int mxo_put_1d_array_on_cart(Ctx* c, int lev, const mx_basestate* s0, mx_multifab* s0_cart,
                             int is_input_edge_centered, int is_output_a_vector,
                             const mx_sph_geom* sg) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    if (!s0 || !s0_cart || s0_cart->nfab != (int)c->grids[lev].size()) return MX_ERR_ARG;
    rc = Put1dArrayOnCart(*c, lev, base(s0), arrs(s0_cart), is_input_edge_centered != 0,
                          is_output_a_vector != 0, sg);
    return finish(rc);
}

int mxo_make_w0mac(Ctx* c, int lev, const mx_basestate* w0, const mx_multifab* w0_cart,
                   mx_face_multifab* w0mac, const mx_sph_geom* sg) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    if (!w0 || !w0mac) return MX_ERR_ARG;
    std::vector<Arr> wc;
    if (sg && sg->w0mac_interp_type == 1) {
        if (!w0_cart || w0_cart->nfab != (int)c->grids[lev].size()) return MX_ERR_ARG;
        wc = arrs(w0_cart);
    }
    Face wm = faces(w0mac);
    rc = MakeW0mac(*c, lev, base(w0), wc, wm.d, sg);
    return finish(rc);
}

int mxo_make_s0mac(Ctx* c, int lev, const mx_basestate* s0, const mx_multifab* s0_cart,
                   mx_face_multifab* s0mac, const mx_sph_geom* sg) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    if (!s0 || !s0mac) return MX_ERR_ARG;
    std::vector<Arr> sc;
    if (sg && sg->s0mac_interp_type == 1) {
        if (!s0_cart || s0_cart->nfab != (int)c->grids[lev].size()) return MX_ERR_ARG;
        sc = arrs(s0_cart);
    }
    Face sm = faces(s0mac);
    rc = MakeS0mac(*c, lev, base(s0), sc, sm.d, sg);
    return finish(rc);
}

int mxo_put_in_pert_form(Ctx* c, int lev, mx_multifab* scal, const mx_basestate* s0, int comp,
                         int flag, const mx_sph_geom* sg) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    if (!scal || !s0 || comp < 0 || comp >= c->p.nscal) return MX_ERR_ARG;
    rc = PutInPertForm(*c, lev, arrs(scal), base(s0), comp, flag != 0, sg);
    return finish(rc);
}

int mxo_make_vel_force(Ctx* c, int lev, mx_multifab* vel_force, const mx_face_multifab* uedge,
                       const mx_multifab* rho, const mx_multifab* gpi,
                       const mx_multifab* w0_force_cart, const mx_multifab* w0_cart,
                       const mx_multifab* normal, const mx_basestate* rho0,
                       const mx_basestate* grav_cell, const mx_basestate* w0,
                       double buoyancy_cutoff_factor, double base_cutoff_density,
                       int do_add_utilde_force, int average_base_state, const mx_sph_geom* sg) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    const size_t nb = c->grids[lev].size();
    auto ok = [&](const mx_multifab* m) { return m && m->nfab == (int)nb; };
    if (!ok(vel_force) || !ok(rho) || !ok(gpi) || !ok(w0_force_cart) || !rho0 || !grav_cell)
        return MX_ERR_ARG;
    const bool sph = c->p.spherical != 0;
    if (sph && (!ok(normal) || !w0)) return MX_ERR_ARG;
    if (!sph && do_add_utilde_force && !ok(w0_cart)) return MX_ERR_ARG;
    if (do_add_utilde_force && !uedge) return MX_ERR_ARG;
    if (do_add_utilde_force != 0 && do_add_utilde_force != 1) return MX_ERR_ARG;
    Face ue = faces(uedge);
    std::vector<Arr> w0c = (!sph && do_add_utilde_force) ? arrs(w0_cart) : std::vector<Arr>();
    std::vector<Arr> nrm = sph ? arrs(normal) : std::vector<Arr>();
    rc = MakeVelForce(*c, lev, arrs(vel_force), ue.d, arrs(rho), arrs(gpi), arrs(w0_force_cart),
                      w0c, nrm, base(rho0), base(grav_cell), base(w0), buoyancy_cutoff_factor,
                      base_cutoff_density, do_add_utilde_force != 0, average_base_state != 0, sg);
    return finish(rc);
}

int mxo_modify_scal_force(Ctx* c, int lev, mx_multifab* scal_force, const mx_multifab* state,
                          const mx_face_multifab* umac, const mx_basestate* s0_edge,
                          const mx_multifab* s0_cart, const mx_multifab* w0_cart,
                          const mx_basestate* w0, int comp, int fullform, const mx_sph_geom* sg) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    const size_t nb = c->grids[lev].size();
    auto ok = [&](const mx_multifab* m) { return m && m->nfab == (int)nb; };
    if (!ok(scal_force) || !ok(state) || !umac || comp < 0 || comp >= c->p.nscal)
        return MX_ERR_ARG;
    const bool sph = c->p.spherical != 0;
    if (sph && !w0) return MX_ERR_ARG;
    if (!sph && !ok(w0_cart)) return MX_ERR_ARG;
    if (!fullform && (!ok(s0_cart) || (!sph && !s0_edge))) return MX_ERR_ARG;
    Face um = faces(umac);
    std::vector<Arr> s0c = fullform ? std::vector<Arr>() : arrs(s0_cart);
    std::vector<Arr> w0c = sph ? std::vector<Arr>() : arrs(w0_cart);
    rc = ModifyScalForce(*c, lev, arrs(scal_force), arrs(state), um.d, base(s0_edge), s0c, w0c,
                         base(w0), comp, fullform != 0, sg);
    return finish(rc);
}

int mxo_average_down(Ctx* c, int lev, const mx_multifab* fine, mx_multifab* crse, int comp,
                     int ncomp) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    if (lev + 1 > c->p.finest_level || c->grids[lev + 1].empty() || !fine || !crse ||
        fine->nfab != (int)c->grids[lev + 1].size() || crse->nfab != (int)c->grids[lev].size())
        return MX_ERR_ARG;
    return finish(AverageDown(*c, lev, arrs(fine), arrs(crse), comp, ncomp));
}

int mxo_average_down_faces(Ctx* c, int lev, const mx_face_multifab* fine,
                           mx_face_multifab* crse) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    if (lev + 1 > c->p.finest_level || c->grids[lev + 1].empty() || !fine || !crse)
        return MX_ERR_ARG;
    Face f = faces(fine), cr = faces(crse);
    for (int d = 0; d < c->p.dim; ++d)
        if (f.d[d].size() != c->grids[lev + 1].size() || cr.d[d].size() != c->grids[lev].size())
            return MX_ERR_ARG;
    return finish(AverageDownFaces(*c, lev, f.d, cr.d));
}

int mxo_fill_patch(Ctx* c, int lev, mx_multifab* mf, const mx_multifab* src, int srccomp,
                   int destcomp, int ncomp, const mx_bcrec* bcs, int variable_type) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    if (lev != 0) return MX_ERR_UNSUPPORTED;  // FillPatchTwoLevels is not restated
    const size_t nb = c->grids[lev].size();
    if (!mf || !src || !bcs || mf->nfab != (int)nb || src->nfab != (int)nb || srccomp < 0 ||
        destcomp < 0 || ncomp < 1)
        return MX_ERR_ARG;
    std::vector<Arr> m = arrs(mf), s = arrs(src);
    if (nb > 0 && m[0].p == s[0].p && srccomp != destcomp && srccomp < destcomp + ncomp &&
        destcomp < srccomp + ncomp)
        return MX_ERR_ARG;  // overlapping component ranges of one MultiFab
    for (size_t b = 0; b < nb; ++b) {
        const IBox& g = c->grids[lev][b];
        if (!m[b].valid() || !s[b].valid() || m[b].nc < destcomp + ncomp ||
            s[b].nc < srccomp + ncomp)
            return MX_ERR_ARG;
        for (int d = 0; d < 3; ++d) {
            const int ng = g.lo[d] - m[b].b[d];
            if (ng < 0 || m[b].e[d] - 1 - g.hi[d] != ng || s[b].b[d] > g.lo[d] ||
                s[b].e[d] - 1 < g.hi[d])
                return MX_ERR_BOX;
            if (b > 0 && ng != c->grids[lev][0].lo[d] - m[0].b[d]) return MX_ERR_BOX;
        }
    }
    return finish(FillPatchSingleLevel(*c, lev, m, s, srccomp, destcomp, ncomp, bcs,
                                       variable_type == 1));
}

int mxo_convert_rhox_to_x(Ctx* c, int lev, mx_multifab* scal, int flag) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    if (!scal) return MX_ERR_ARG;
    return finish(ConvertByRho(*c, lev, arrs(scal), c->p.first_spec, c->p.num_spec, flag != 0));
}

int mxo_convert_rhoh_to_h(Ctx* c, int lev, mx_multifab* scal, int flag) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    if (!scal) return MX_ERR_ARG;
    return finish(ConvertByRho(*c, lev, arrs(scal), c->p.comp_rhoh, 1, flag != 0));
}

// nanoseconds spent in temporary allocation since the last reset
long long mxo_alloc_ns(int reset) {
    return reset ? g_alloc_ns.exchange(0) : g_alloc_ns.load();
}

int mxo_omp_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

}  // extern "C"
