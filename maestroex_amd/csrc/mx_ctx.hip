// mx_ctx.hip -- context lifecycle, workspace, argument checks, the ghost-face
// fill (FillBoundary + FillUmacGhost), Addw0, CelltoEdge and the unit-level
// PPM / slope entry points of libmaestroex_amd.so.
#include <algorithm>
#include <array>
#include <cstdio>

#include "mx_internal.H"

using namespace mxd;

namespace mxh {

KGeom kgeom(const mx_adv_ctx* c, int lev) {
    KGeom g;
    g.dim = c->p.dim;
    g.ppm_type = c->p.ppm_type;
    g.slope_order = c->p.slope_order;
    g.ppm_trace_forces = c->p.ppm_trace_forces;
    g.spherical = c->p.spherical;
    g.dt = c->p.dt;
    g.rel_eps = c->p.rel_eps;
    for (int d = 0; d < 3; ++d) {
        g.dx[d] = c->p.dx[lev][d];
        g.domlo[d] = (d < c->p.dim) ? c->p.domain_lo[lev][d] : 0;
        g.domhi[d] = (d < c->p.dim) ? c->p.domain_hi[lev][d] : 0;
        g.phys_lo[d] = c->p.phys_bc_lo[d];
        g.phys_hi[d] = c->p.phys_bc_hi[d];
    }
    g.xcd = c->xcd_remap;
    return g;
}

int ws_reserve(mx_adv_ctx* c, size_t bytes) {
    c->ws_off = 0;
    if (c->dry) {
        c->dry_ws = std::max(c->dry_ws, bytes);
        return MX_OK;
    }
    if (c->ext_ws) return bytes <= c->ext_ws_cap ? MX_OK : MX_ERR_NOMEM;
    if (bytes <= c->ws_cap) return MX_OK;
    MX_HIP_CHECK(hipStreamSynchronize(c->stream));
    if (c->ws) MX_HIP_CHECK(hipFree(c->ws));
    c->ws = nullptr;
    c->ws_cap = 0;
    size_t cap = bytes + bytes / 8 + (1 << 20);
    if (hipMalloc((void**)&c->ws, cap) != hipSuccess) {
        c->ws = nullptr;
        return MX_ERR_NOMEM;
    }
    c->ws_cap = cap;
    return MX_OK;
}

hipEvent_t prof_event(mx_adv_ctx* c) {
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_used++];
}

double* ws_take(mx_adv_ctx* c, size_t ndouble) {
    size_t off = (c->ws_off + 255) & ~(size_t)255;
    c->ws_off = off + ndouble * sizeof(double);
    // a dry run only sizes the workspace: its views are never dereferenced
    const uintptr_t base = c->dry ? 0 : reinterpret_cast<uintptr_t>(c->ext_ws ? c->ext_ws : c->ws);
    return reinterpret_cast<double*>(base + off);
}

int Table::upload() {
    size_t bytes = c->htab.size();
    if (c->dry) {
        c->dry_tab += (bytes + 255) & ~(size_t)255;
        base = nullptr;
        return MX_OK;
    }
    if (c->ext_tab) {
        // caller-given device memory: one region per table of the call, in
        // order; the host bytes are kept until the stream has copied them
        const size_t off = (c->ext_tab_off + 255) & ~(size_t)255;
        if (off + bytes > c->ext_tab_cap) return MX_ERR_NOMEM;
        base = c->ext_tab + off;
        c->ext_tab_off = off + bytes;
        if (bytes == 0) return MX_OK;
        c->ext_host.push_back(c->htab);
        MX_HIP_CHECK(hipMemcpyAsync(base, c->ext_host.back().data(), bytes, hipMemcpyHostToDevice,
                                    c->stream));
        return MX_OK;
    }
    if (c->tab_mode == 2) {
        // replay under capture: the recorded table of this position (the
        // same call sequence over the same descriptors; the bytes are not
        // compared, the host structs carry uninitialised padding)
        if (c->rec_next >= c->rec_tabs.size()) return MX_ERR_ARG;
        if (c->rec_tabs[c->rec_next].size() != bytes) return MX_ERR_ARG;
        base = c->rec_arena + c->rec_offs[c->rec_next++];
        return MX_OK;
    }
    if (c->tab_mode == 1) c->rec_tabs.push_back(c->htab);
    if (bytes == 0) {
        base = nullptr;
        return MX_OK;
    }
    // an FNV-1a-style hash over 8-byte words (tables of many thousand
    // entries are hashed on every call), then a full comparison
    unsigned long long h = 1469598103934665603ULL;
    size_t i = 0;
    for (; i + 8 <= bytes; i += 8) {
        unsigned long long w;
        std::memcpy(&w, c->htab.data() + i, 8);
        h = (h ^ w) * 1099511628211ULL;
        h ^= h >> 29;
    }
    for (; i < bytes; ++i) h = (h ^ (unsigned char)c->htab[i]) * 1099511628211ULL;
    for (const auto& e : c->tab_cache) {
        if (e.hash == h && e.bytes.size() == bytes &&
            std::memcmp(e.bytes.data(), c->htab.data(), bytes) == 0) {
            base = e.dev;
            return MX_OK;
        }
    }
    constexpr size_t kCacheEntries = 4096, kCacheBytes = (size_t)256 << 20;
    if (c->tab_cache.size() >= kCacheEntries || c->tab_cache_bytes + bytes > kCacheBytes) {
        // full: queued kernels may still read the cached copies
        MX_HIP_CHECK(hipStreamSynchronize(c->stream));
        free_tab_cache(c);
    }
    char* dev = nullptr;
    if (hipMalloc((void**)&dev, bytes) != hipSuccess) return MX_ERR_NOMEM;
    c->tab_cache.push_back({h, c->htab, dev});
    c->tab_cache_bytes += bytes;
    if (c->tab_slots.empty()) c->tab_slots.resize(mx_adv_ctx::kTabSlots);
    mx_adv_ctx::TabSlot& t = c->tab_slots[c->tab_next];
    c->tab_next = (c->tab_next + 1) % mx_adv_ctx::kTabSlots;
    if (t.pending) {
        MX_HIP_CHECK(hipEventSynchronize(t.done));
        t.pending = false;
    }
    if (!t.done) MX_HIP_CHECK(hipEventCreateWithFlags(&t.done, hipEventDisableTiming));
    if (bytes > t.cap) {
        free_tab_slot(t);
        size_t cap = std::max(bytes * 2, (size_t)1 << 16);
        if (hipHostMalloc((void**)&t.host, cap, hipHostMallocDefault) != hipSuccess) {
            t.host = nullptr;
            return MX_ERR_NOMEM;
        }
        t.cap = cap;
    }
    std::memcpy(t.host, c->htab.data(), bytes);
    MX_HIP_CHECK(hipMemcpyAsync(dev, t.host, bytes, hipMemcpyHostToDevice, c->stream));
    MX_HIP_CHECK(hipEventRecord(t.done, c->stream));
    t.pending = true;
    base = dev;
    return MX_OK;
}

void free_tab_cache(mx_adv_ctx* c) {
    for (auto& e : c->tab_cache) (void)hipFree(e.dev);
    c->tab_cache.clear();
    c->tab_cache_bytes = 0;
}

void free_tab_slot(mx_adv_ctx::TabSlot& t) {
    if (t.host) (void)hipHostFree(t.host);
    t.host = nullptr;
    t.cap = 0;
}

int check_level(const mx_adv_ctx* c, int lev) {
    if (!c || lev < 0 || lev > c->p.finest_level) return MX_ERR_ARG;
    if (c->grids[lev].empty()) return MX_ERR_ARG;
    return MX_OK;
}

int check_mf(const mx_adv_ctx* c, int lev, const mx_multifab* m, int ncomp, int ngrow,
             int nodal_dir) {
    if (!m || !m->fab) return MX_ERR_ARG;
    const auto& g = c->grids[lev];
    if (m->nfab != (int)g.size()) return MX_ERR_ARG;
    for (size_t b = 0; b < g.size(); ++b) {
        HBox r = hbox(g[b]);
        if (nodal_dir >= 0) r = nodal(r, nodal_dir);
        r = grow(r, c->p.dim, ngrow);
        if (!covers(m->fab[b], r, ncomp)) return MX_ERR_BOX;
    }
    return MX_OK;
}

int check_face(const mx_adv_ctx* c, int lev, const mx_face_multifab* f, int ncomp, int ngrow) {
    if (!f) return MX_ERR_ARG;
    for (int d = 0; d < c->p.dim; ++d) {
        int rc = check_mf(c, lev, &f->dir[d], ncomp, ngrow, d);
        if (rc) return rc;
    }
    return MX_OK;
}

}  // namespace mxh

using namespace mxh;

// ---------------------------------------------------------------------------
// FillBoundary (single level, periodic) of a face MultiFab

namespace {

struct SlabDesc {
    DA dst;
    int lo[3], hi[3];
    int cand_off, cand_n;
    int ncomp;
    int vlo[3], vhi[3];  // valid faces of the destination box
    int plo[3], phi[3];  // faces not outside a non-periodic domain side
    int pad_ = 0;
};

struct CandDesc {
    DA src;
    int vlo[3], vhi[3];  // valid faces of the source box
    int shift[3];        // source index = destination index - shift
    int pad_ = 0;
};

// one ghost face: copy from the first candidate source that holds it
__global__ void k_fill_ghost_faces(const SlabDesc* __restrict__ slabs,
                                   const CandDesc* __restrict__ cands) {
    const SlabDesc& t = slabs[blockIdx.y];
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!region_point(idx, t.lo, t.hi, i, j, k)) return;
    for (int q = 0; q < t.cand_n; ++q) {
        const CandDesc& cd = cands[t.cand_off + q];
        const int si = i - cd.shift[0], sj = j - cd.shift[1], sk = k - cd.shift[2];
        if (si >= cd.vlo[0] && si <= cd.vhi[0] && sj >= cd.vlo[1] && sj <= cd.vhi[1] &&
            sk >= cd.vlo[2] && sk <= cd.vhi[2]) {
            for (int n = 0; n < t.ncomp; ++n) t.dst(i, j, k, n) = cd.src(si, sj, sk, n);
            return;
        }
    }
    // no box of the level covers it: a coarse-fine face of a finer level takes
    // the value of the box's nearest valid face; faces outside a non-periodic
    // domain side are left to FillUmacGhost
    if (i < t.plo[0] || i > t.phi[0] || j < t.plo[1] || j > t.phi[1] || k < t.plo[2] ||
        k > t.phi[2])
        return;
    const int ri = min(max(i, t.vlo[0]), t.vhi[0]);
    const int rj = min(max(j, t.vlo[1]), t.vhi[1]);
    const int rk = min(max(k, t.vlo[2]), t.vhi[2]);
    for (int n = 0; n < t.ncomp; ++n) t.dst(i, j, k, n) = t.dst(ri, rj, rk, n);
}

// FillUmacGhost (MaestroFillData.cpp:176-453) for faces normal to D; the
// kernel covers the ghost plane at domlo-1 (side 0) or domhi+2 (side 1) of
// one box's grown face box
struct UGTask {
    DA u, v, w;
    int lo[3], hi[3];
    int side;
    int bc;
};

template <int D>
__global__ void k_fill_umac_ghost(const UGTask* __restrict__ tasks, int dim) {
    const UGTask& t = tasks[blockIdx.y];
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!region_point(idx, t.lo, t.hi, i, j, k)) return;
    const DA& umac = t.u;
    const DA& vmac = t.v;
    const DA& wmac = t.w;
    const bool three = (dim == 3);
    if (D == 0) {
        if (t.side == 0) {
            switch (t.bc) {
                case MX_PHYS_INFLOW:
                    umac(i, j, k) = umac(i + 1, j, k);
                    vmac(i, j, k) = 0.0;
                    if (three) wmac(i, j, k) = 0.0;
                    break;
                case MX_PHYS_SLIPWALL:
                case MX_PHYS_NOSLIPWALL:
                    umac(i, j, k) = 0.0;
                    vmac(i, j, k) = 0.0;
                    if (three) wmac(i, j, k) = 0.0;
                    break;
                case MX_PHYS_OUTFLOW:
                    umac(i, j, k) = umac(i + 1, j, k);
                    vmac(i, j, k) = vmac(i + 1, j, k);
                    if (three) wmac(i, j, k) = wmac(i + 1, j, k);
                    break;
                case MX_PHYS_SYMMETRY:
                    umac(i, j, k) = -umac(i + 2, j, k);
                    vmac(i, j, k) = vmac(i + 1, j, k);
                    if (three) wmac(i, j, k) = wmac(i + 1, j, k);
                    break;
                default:
                    break;
            }
        } else {
            switch (t.bc) {
                case MX_PHYS_INFLOW:
                    umac(i, j, k) = umac(i - 1, j, k);
                    vmac(i - 1, j, k) = 0.0;
                    if (three) wmac(i - 1, j, k) = 0.0;
                    break;
                case MX_PHYS_SLIPWALL:
                case MX_PHYS_NOSLIPWALL:
                    umac(i, j, k) = 0.0;
                    vmac(i - 1, j, k) = 0.0;
                    if (three) wmac(i - 1, j, k) = 0.0;
                    break;
                case MX_PHYS_OUTFLOW:
                    umac(i, j, k) = umac(i - 1, j, k);
                    vmac(i - 1, j, k) = vmac(i - 2, j, k);
                    if (three) wmac(i - 1, j, k) = wmac(i - 2, j, k);
                    break;
                case MX_PHYS_SYMMETRY:
                    umac(i, j, k) = -umac(i - 2, j, k);
                    vmac(i - 1, j, k) = vmac(i - 2, j, k);
                    if (three) wmac(i - 1, j, k) = wmac(i - 2, j, k);
                    break;
                default:
                    break;
            }
        }
    } else if (D == 1) {
        if (t.side == 0) {
            switch (t.bc) {
                case MX_PHYS_INFLOW:
                    umac(i, j, k) = 0.0;
                    vmac(i, j, k) = vmac(i, j + 1, k);
                    if (three) wmac(i, j, k) = 0.0;
                    break;
                case MX_PHYS_SLIPWALL:
                case MX_PHYS_NOSLIPWALL:
                    umac(i, j, k) = 0.0;
                    vmac(i, j, k) = 0.0;
                    if (three) wmac(i, j, k) = 0.0;
                    break;
                case MX_PHYS_OUTFLOW:
                    umac(i, j, k) = umac(i, j + 1, k);
                    vmac(i, j, k) = vmac(i, j + 1, k);
                    if (three) wmac(i, j, k) = wmac(i, j + 1, k);
                    break;
                case MX_PHYS_SYMMETRY:
                    umac(i, j, k) = umac(i, j + 1, k);
                    vmac(i, j, k) = -vmac(i, j + 2, k);
                    if (three) wmac(i, j, k) = wmac(i, j + 1, k);
                    break;
                default:
                    break;
            }
        } else {
            switch (t.bc) {
                case MX_PHYS_INFLOW:
                    umac(i, j - 1, k) = 0.0;
                    vmac(i, j, k) = vmac(i, j - 1, k);
                    if (three) wmac(i, j - 1, k) = 0.0;
                    break;
                case MX_PHYS_SLIPWALL:
                case MX_PHYS_NOSLIPWALL:
                    umac(i, j - 1, k) = 0.0;
                    vmac(i, j, k) = 0.0;
                    if (three) wmac(i, j - 1, k) = 0.0;
                    break;
                case MX_PHYS_OUTFLOW:
                    umac(i, j - 1, k) = umac(i, j - 2, k);
                    vmac(i, j, k) = vmac(i, j - 1, k);
                    if (three) wmac(i, j - 1, k) = wmac(i, j - 2, k);
                    break;
                case MX_PHYS_SYMMETRY:
                    umac(i, j - 1, k) = umac(i, j - 2, k);
                    vmac(i, j, k) = -vmac(i, j - 2, k);
                    if (three) wmac(i, j - 1, k) = wmac(i, j - 2, k);
                    break;
                default:
                    break;
            }
        }
    } else {
        if (t.side == 0) {
            switch (t.bc) {
                case MX_PHYS_INFLOW:
                    umac(i, j, k) = 0.0;
                    vmac(i, j, k) = 0.0;
                    wmac(i, j, k) = wmac(i, j, k + 1);
                    break;
                case MX_PHYS_SLIPWALL:
                case MX_PHYS_NOSLIPWALL:
                    umac(i, j, k) = 0.0;
                    vmac(i, j, k) = 0.0;
                    wmac(i, j, k) = 0.0;
                    break;
                case MX_PHYS_OUTFLOW:
                    umac(i, j, k) = umac(i, j, k + 1);
                    vmac(i, j, k) = vmac(i, j, k + 1);
                    wmac(i, j, k) = wmac(i, j, k + 1);
                    break;
                case MX_PHYS_SYMMETRY:
                    umac(i, j, k) = umac(i, j, k + 1);
                    vmac(i, j, k) = vmac(i, j, k + 1);
                    wmac(i, j, k) = -wmac(i, j, k + 2);
                    break;
                default:
                    break;
            }
        } else {
            switch (t.bc) {
                case MX_PHYS_INFLOW:
                    umac(i, j, k - 1) = 0.0;
                    vmac(i, j, k - 1) = 0.0;
                    wmac(i, j, k) = wmac(i, j, k - 1);
                    break;
                case MX_PHYS_SLIPWALL:
                case MX_PHYS_NOSLIPWALL:
                    umac(i, j, k - 1) = 0.0;
                    vmac(i, j, k - 1) = 0.0;
                    wmac(i, j, k) = 0.0;
                    break;
                case MX_PHYS_OUTFLOW:
                    umac(i, j, k - 1) = umac(i, j, k - 2);
                    vmac(i, j, k - 1) = vmac(i, j, k - 2);
                    wmac(i, j, k) = wmac(i, j, k - 1);
                    break;
                case MX_PHYS_SYMMETRY:
                    umac(i, j, k - 1) = umac(i, j, k - 2);
                    vmac(i, j, k - 1) = vmac(i, j, k - 2);
                    wmac(i, j, k) = -wmac(i, j, k - 2);
                    break;
                default:
                    break;
            }
        }
    }
}

const mx_adv_ctx::GhostPlan& ghost_plan(mx_adv_ctx* c, int lev, int ndir, const mx_multifab& m) {
    const int dim = c->p.dim;
    const auto& grids = c->grids[lev];
    HBox v0 = nodal(hbox(grids[0]), ndir);
    long long ng = v0.lo[0] - m.fab[0].begin[0];
    auto key = std::make_tuple(lev, ndir, ng);
    auto it = c->plans.find(key);
    if (it != c->plans.end()) return it->second;

    mx_adv_ctx::GhostPlan plan;
    int nlen[3];
    for (int d = 0; d < 3; ++d) nlen[d] = c->p.domain_hi[lev][d] - c->p.domain_lo[lev][d] + 1;
    std::vector<std::array<int, 3>> shifts;
    for (int sx = -1; sx <= 1; ++sx)
        for (int sy = -1; sy <= 1; ++sy)
            for (int sz = -1; sz <= 1; ++sz) {
                int s[3] = {sx, sy, sz};
                bool ok = true;
                for (int d = 0; d < 3; ++d)
                    if (s[d] != 0 && (d >= dim || !c->p.is_periodic[d])) ok = false;
                if (ok) shifts.push_back({sx * nlen[0], sy * nlen[1], sz * nlen[2]});
            }
    for (size_t b = 0; b < grids.size(); ++b) {
        HBox valid = nodal(hbox(grids[b]), ndir);
        HBox fb = grow(valid, dim, (int)ng);
        const int off = (int)plan.cands.size();
        for (size_t s = 0; s < grids.size(); ++s) {
            HBox sv = nodal(hbox(grids[s]), ndir);
            for (const auto& sh : shifts) {
                bool empty = false;
                for (int d = 0; d < 3; ++d) {
                    if (std::max(fb.lo[d], sv.lo[d] + sh[d]) > std::min(fb.hi[d], sv.hi[d] + sh[d]))
                        empty = true;
                }
                if (empty) continue;
                if (s == b && sh[0] == 0 && sh[1] == 0 && sh[2] == 0) continue;
                plan.cands.push_back({(int)s, {sh[0], sh[1], sh[2]}});
            }
        }
        const int n = (int)plan.cands.size() - off;
        // the ghost shell fb \ valid as up to 2*dim slabs
        for (int d = 0; d < dim; ++d) {
            for (int side = 0; side < 2; ++side) {
                mx_adv_ctx::GhostSlab sl;
                sl.box = (int)b;
                for (int q = 0; q < 3; ++q) {
                    if (q < d) {
                        sl.lo[q] = valid.lo[q];
                        sl.hi[q] = valid.hi[q];
                    } else {
                        sl.lo[q] = fb.lo[q];
                        sl.hi[q] = fb.hi[q];
                    }
                }
                if (side == 0) {
                    sl.hi[d] = valid.lo[d] - 1;
                } else {
                    sl.lo[d] = valid.hi[d] + 1;
                }
                if (sl.lo[d] > sl.hi[d]) continue;
                sl.cand_off = off;
                sl.cand_n = n;
                plan.slabs.push_back(sl);
            }
        }
    }
    return c->plans.emplace(key, std::move(plan)).first->second;
}

}  // namespace

namespace mxh {

int fill_face_ghosts(mx_adv_ctx* c, int lev, const mx_face_multifab* uedge, bool extrapolate) {
    const int dim = c->p.dim;
    const auto& grids = c->grids[lev];
    const int block = 256;
    // FillBoundary: every direction's ghost faces are filled from that
    // direction's valid faces, so the slabs of all directions go in one
    // launch
    std::vector<SlabDesc> slabs;
    std::vector<CandDesc> cands;
    long long maxpts = 0;
    for (int d = 0; d < dim; ++d) {
        const mx_multifab& m = uedge->dir[d];
        const auto& plan = ghost_plan(c, lev, d, m);
        if (plan.slabs.empty()) continue;
        const int cbase = (int)cands.size();
        for (const auto& q : plan.cands) {
            CandDesc cd;
            cd.src = to_da(m.fab[q.src_box]);
            HBox v = nodal(hbox(grids[q.src_box]), d);
            for (int r = 0; r < 3; ++r) {
                cd.vlo[r] = v.lo[r];
                cd.vhi[r] = v.hi[r];
                cd.shift[r] = q.shift[r];
            }
            cands.push_back(cd);
        }
        for (const auto& sl : plan.slabs) {
            SlabDesc sd;
            sd.dst = to_da(m.fab[sl.box]);
            HBox r;
            for (int q = 0; q < 3; ++q) {
                sd.lo[q] = r.lo[q] = sl.lo[q];
                sd.hi[q] = r.hi[q] = sl.hi[q];
            }
            sd.cand_off = cbase + sl.cand_off;
            sd.cand_n = sl.cand_n;
            sd.ncomp = m.fab[sl.box].ncomp;
            HBox vb = nodal(hbox(grids[sl.box]), d);
            for (int q = 0; q < 3; ++q) {
                sd.vlo[q] = vb.lo[q];
                sd.vhi[q] = vb.hi[q];
                const bool bounded = q < dim && !c->p.is_periodic[q];
                sd.plo[q] = bounded ? c->p.domain_lo[lev][q] : -(1 << 30);
                sd.phi[q] = bounded ? c->p.domain_hi[lev][q] + (q == d ? 1 : 0) : (1 << 30);
                if (!extrapolate) {
                    // an empty region: every uncovered face is left as it is
                    sd.plo[q] = 1 << 30;
                    sd.phi[q] = -(1 << 30);
                }
            }
            maxpts = std::max(maxpts, npts(r));
            slabs.push_back(sd);
        }
    }
    if (!slabs.empty()) {
        Table tab(c);
        size_t os = tab.add(slabs);
        size_t oc = tab.add(cands);
        int rc = tab.upload();
        if (rc) return rc;
        // grid y takes the slabs, at most 65535 per launch
        const SlabDesc* ds = tab.dev<SlabDesc>(os);
        for (size_t s0 = 0; s0 < slabs.size(); s0 += 65535) {
            const int ns = (int)std::min<size_t>(65535, slabs.size() - s0);
            MX_LAUNCH(c, "k_fill_ghost_faces", k_fill_ghost_faces,
                      grid_for(maxpts, ns, 1, block), dim3(block), ds + s0,
                      tab.dev<CandDesc>(oc));
        }
    }
    // FillUmacGhost: x, then y, then z (each reads what the previous wrote)
    for (int d = 0; d < dim; ++d) {
        std::vector<UGTask> tasks;
        long long maxpts = 0;
        for (int side = 0; side < 2; ++side) {
            const int bc = side == 0 ? c->p.phys_bc_lo[d] : c->p.phys_bc_hi[d];
            if (bc == MX_PHYS_INTERIOR) continue;
            const int plane = side == 0 ? c->p.domain_lo[lev][d] - 1 : c->p.domain_hi[lev][d] + 2;
            for (size_t b = 0; b < grids.size(); ++b) {
                HBox r = grow(nodal(hbox(grids[b]), d), dim, 1);
                if (plane < r.lo[d] || plane > r.hi[d]) continue;
                r.lo[d] = r.hi[d] = plane;
                UGTask t;
                t.u = to_da(uedge->dir[0].fab[b]);
                t.v = to_da(uedge->dir[1].fab[b]);
                t.w = dim == 3 ? to_da(uedge->dir[2].fab[b]) : null_da();
                for (int q = 0; q < 3; ++q) {
                    t.lo[q] = r.lo[q];
                    t.hi[q] = r.hi[q];
                }
                t.side = side;
                t.bc = bc;
                maxpts = std::max(maxpts, npts(r));
                tasks.push_back(t);
            }
        }
        if (tasks.empty()) continue;
        Table tab(c);
        size_t off = tab.add(tasks);
        int rc = tab.upload();
        if (rc) return rc;
        dim3 gr = grid_for(maxpts, (int)tasks.size(), 1, block);
        if (d == 0)
            MX_LAUNCH(c, "k_fill_umac_ghost", k_fill_umac_ghost<0>,  gr, dim3(block), tab.dev<UGTask>(off), dim);
        else if (d == 1)
            MX_LAUNCH(c, "k_fill_umac_ghost", k_fill_umac_ghost<1>,  gr, dim3(block), tab.dev<UGTask>(off), dim);
        else
            MX_LAUNCH(c, "k_fill_umac_ghost", k_fill_umac_ghost<2>,  gr, dim3(block), tab.dev<UGTask>(off), dim);

    }
    return MX_OK;
}

}  // namespace mxh

// ---------------------------------------------------------------------------
// Addw0 (MaestroFill3dData.cpp:357-461)

namespace {

struct AddTask {
    DA e;
    DA wm;  // spherical: w0mac of this direction
    int lo[3], hi[3];
    int r_dir;  // planar: radial index direction
    int pad_ = 0;
};

__global__ void k_addw0(const AddTask* __restrict__ tasks, const double* __restrict__ w0,
                        long long w0_off, double mult, int spherical) {
    const AddTask& t = tasks[blockIdx.y];
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!region_point(idx, t.lo, t.hi, i, j, k)) return;
    if (!spherical) {
        const int r = t.r_dir == 1 ? j : k;
        t.e(i, j, k) += mult * w0[w0_off + r];
    } else {
        t.e(i, j, k) += mult * t.wm(i, j, k);
    }
}

}  // namespace

namespace mxh {

int addw0(mx_adv_ctx* c, int lev, const mx_face_multifab* uedge, const mx_face_multifab* w0mac,
          const mx_basestate* w0, double mult, bool fill) {
    const int dim = c->p.dim;
    const auto& grids = c->grids[lev];
    std::vector<AddTask> tasks;
    long long maxpts = 0;
    if (!c->p.spherical) {
        if (!w0 || !w0->p) return MX_ERR_ARG;
        const int d = dim - 1;
        for (size_t b = 0; b < grids.size(); ++b) {
            HBox r = nodal(hbox(grids[b]), d);
            for (int t = 0; t < dim; ++t)
                if (t != d) r = grow_dir(r, t, 1);
            AddTask a;
            a.e = to_da(uedge->dir[d].fab[b]);
            a.wm = null_da();
            for (int q = 0; q < 3; ++q) {
                a.lo[q] = r.lo[q];
                a.hi[q] = r.hi[q];
            }
            a.r_dir = d;
            maxpts = std::max(maxpts, npts(r));
            tasks.push_back(a);
        }
    } else {
        if (dim != 3) return MX_ERR_UNSUPPORTED;
        if (!w0mac) return MX_ERR_ARG;
        for (int d = 0; d < 3; ++d)
            for (size_t b = 0; b < grids.size(); ++b) {
                HBox r = nodal(hbox(grids[b]), d);
                AddTask a;
                a.e = to_da(uedge->dir[d].fab[b]);
                a.wm = to_da(w0mac->dir[d].fab[b]);
                for (int q = 0; q < 3; ++q) {
                    a.lo[q] = r.lo[q];
                    a.hi[q] = r.hi[q];
                }
                a.r_dir = 2;
                maxpts = std::max(maxpts, npts(r));
                tasks.push_back(a);
            }
    }
    Table tab(c);
    size_t off = tab.add(tasks);
    int rc = tab.upload();
    if (rc) return rc;
    const int block = 256;
    MX_LAUNCH(c, "k_addw0", k_addw0,  grid_for(maxpts, (int)tasks.size(), 1, block), dim3(block), tab.dev<AddTask>(off), c->p.spherical ? nullptr : w0->p,
                       c->p.spherical ? 0LL : (long long)lev * w0->len, mult, c->p.spherical);

    return fill ? fill_face_ghosts(c, lev, uedge) : MX_OK;
}

}  // namespace mxh

// ---------------------------------------------------------------------------
// CelltoEdge (MaestroCelltoEdge.cpp:5-55)

namespace {

__global__ void k_cell_to_edge(const double* __restrict__ cell, double* __restrict__ edge,
                               int nr) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > nr) return;
    const int lo = 0, hi = nr;
    const double nr_lev = nr;
    double v;
    if (r == 0) {
        v = cell[r];
    } else if (r == 1 || r == lo) {
        v = 0.5 * (cell[r - 1] + cell[r]);
    } else if (r == nr_lev) {
        v = cell[r - 1];
    } else if (r == nr_lev - 1 || r == hi) {
        v = 0.5 * (cell[r] + cell[r - 1]);
    } else {
        double tmp = 7.0 / 12.0 * (cell[r] + cell[r - 1]) - 1.0 / 12.0 * (cell[r + 1] + cell[r - 2]);
        double s0min = dmin(cell[r], cell[r - 1]);
        double s0max = dmax(cell[r], cell[r - 1]);
        v = dmin(dmax(tmp, s0min), s0max);
    }
    edge[r] = v;
}

// ---------------------------------------------------------------------------
// unit-level PPM and slopes

struct PpmUnitArgs {
    DA s, u, v, w, Ip, Im;
    int lo[3], hi[3];
    int bclo[3], bchi[3];
    int comp;
    int is_umac;
};

template <int PPM, int D>
__device__ __forceinline__ void ppm_unit_dir(const PpmUnitArgs& a, const KGeom& g, int i, int j,
                                             int k) {
    const DA& vel = D == 0 ? a.u : (D == 1 ? a.v : a.w);
    const long long st = a.s.stride(D);
    const double* q = a.s.p + a.s.off(i, j, k, a.comp);
    double v[9];
    load_stencil<PPM>(q, st, v);
    const int pos = D == 0 ? i : (D == 1 ? j : k);
    double sm, sp, Ip, Im;
    ppm_edges<PPM, D>(v, pos, g.domlo[D], g.domhi[D], a.bclo[D], a.bchi[D], sm, sp);
    double uhi, ulo;
    if (a.is_umac) {
        uhi = vel(i + (D == 0), j + (D == 1), k + (D == 2));
        ulo = vel(i, j, k);
    } else {
        uhi = ulo = vel(i, j, k);
    }
    ppm_trace(v[4], sm, sp, uhi, ulo, g.dt, g.dx[D], g.rel_eps, Ip, Im);
    a.Ip(i, j, k, D) = Ip;
    a.Im(i, j, k, D) = Im;
}

template <int PPM>
__global__ void k_ppm_unit(PpmUnitArgs a, KGeom g) {
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!region_point(idx, a.lo, a.hi, i, j, k)) return;
    const int d = blockIdx.z;
    if (d == 0)
        ppm_unit_dir<PPM, 0>(a, g, i, j, k);
    else if (d == 1)
        ppm_unit_dir<PPM, 1>(a, g, i, j, k);
    else
        ppm_unit_dir<PPM, 2>(a, g, i, j, k);
}

struct SlopeUnitArgs {
    DA s, sl;
    int lo[3], hi[3];
    int bclo[MX_MAX_SCAL], bchi[MX_MAX_SCAL];
    int order;
    int ilo, ihi;
};

template <int D>
__global__ void k_slope_unit(SlopeUnitArgs a) {
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!region_point(idx, a.lo, a.hi, i, j, k)) return;
    const int n = blockIdx.z;
    double v[9];
    load_stencil<1>(a.s.p + a.s.off(i, j, k, n), a.s.stride(D), v);
    const int pos = D == 0 ? i : (D == 1 ? j : k);
    a.sl(i, j, k, n) = slope_point<D>(v, a.order, pos, a.ilo, a.ihi, a.bclo[n], a.bchi[n]);
}

}  // namespace

// ---------------------------------------------------------------------------
// C-ABI: lifecycle and unit-level entry points

extern "C" {

int mx_adv_create(const mx_adv_params* p, mx_adv_ctx** out) {
    if (!p || !out) return MX_ERR_ARG;
    if (p->dim != 2 && p->dim != 3) return MX_ERR_ARG;
    if (p->ppm_type < 0 || p->ppm_type > 2) return MX_ERR_ARG;
    if (p->finest_level < 0 || p->finest_level >= MX_MAX_LEV) return MX_ERR_ARG;
    if (p->nscal <= 0 || p->nscal > MX_MAX_SCAL) return MX_ERR_ARG;
    if (p->spherical && p->dim != 3) return MX_ERR_ARG;
    if (p->reflux_type < 0 || p->reflux_type > 2) return MX_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return MX_ERR_NOLIB;
    mx_adv_ctx* c = new mx_adv_ctx();
    c->p = *p;
    bc_setup_host(c);
    if (hipGetDevice(&c->device) != hipSuccess) {
        delete c;
        return MX_ERR_HIP;
    }
    *out = c;
    return MX_OK;
}

int mx_adv_prof_enable(mx_adv_ctx* c, int on) {
    if (!c) return MX_ERR_ARG;
    MX_HIP_CHECK(hipStreamSynchronize(c->stream));
    c->prof_on = on ? 1 : 0;
    c->prof_recs.clear();
    c->ev_used = 0;
    c->prof_names.clear();
    c->prof_ms.clear();
    c->prof_n.clear();
    return MX_OK;
}

int mx_adv_set_tuning(mx_adv_ctx* c, const char* key, long long value) {
    if (!c || !key) return MX_ERR_ARG;
    const std::string k(key);
    if (k == "edge_scal_fused") {
        if (value != 0 && value != 1) return MX_ERR_ARG;
        c->fused_edge = (int)value;
    } else if (k == "edge_scal_tz") {
        if (value < 1 || value > 4096) return MX_ERR_ARG;
        c->fused_tz = (int)value;
    } else if (k == "edge_scal_tile") {
        if (value < -1 || value > 4) return MX_ERR_ARG;
        c->fused_tile = (int)value;
    } else if (k == "edge_scal_per_box") {
        if (value != 0 && value != 1) return MX_ERR_ARG;
        c->fused_per_box = (int)value;
    } else if (k == "utrans_line") {
        if (value != 0 && value != 1) return MX_ERR_ARG;
        c->utrans_line = (int)value;
    } else if (k == "utrans_chunks") {
        if (value < 1 || value > 64) return MX_ERR_ARG;
        c->utrans_chunks = (int)value;
    } else if (k == "velpred_trans_all") {
        if (value != 0 && value != 1) return MX_ERR_ARG;
        c->vp_trans_all = (int)value;
    } else if (k == "velpred_ppm_all") {
        if (value != 0 && value != 1) return MX_ERR_ARG;
        c->vp_ppm_all = (int)value;
    } else if (k == "velpred_fused") {
        if (value != 0 && value != 1) return MX_ERR_ARG;
        c->vp_fused = (int)value;
    } else if (k == "rhoh_flux_epilogue") {
        if (value != 0 && value != 1) return MX_ERR_ARG;
        c->rhoh_flux_epilogue = (int)value;
    } else if (k == "velpred_fused_comp") {
        if (value != 0 && value != 1) return MX_ERR_ARG;
        c->vp_fused_comp = (int)value;
    } else if (k == "velpred_fused_tile") {
        if (value < 0 || value > 5) return MX_ERR_ARG;
        c->vp_fused_tile = (int)value;
    } else if (k == "velpred_fused_tz") {
        if (value < 1 || value > 4096) return MX_ERR_ARG;
        c->vp_fused_tz = (int)value;
    } else if (k == "xcd_remap") {
        if (value != 0 && value != 1) return MX_ERR_ARG;
        c->xcd_remap = (int)value;
    } else if (k == "adjoint_state_fused") {
        if (value != 0 && value != 1) return MX_ERR_ARG;
        c->adj_state_fused = (int)value;
    } else if (k == "adjoint_face_sums") {
        if (value != 0 && value != 1) return MX_ERR_ARG;
        c->adj_face_sums = (int)value;
    } else if (k == "velpred_fused_xcd") {
        if (value < -1 || value > 2) return MX_ERR_ARG;
        c->vp_fused_xcd = (int)value;
    } else if (k == "edge_scal_xcd") {
        if (value < 0 || value > 2) return MX_ERR_ARG;
        c->es_fused_xcd = (int)value;
    } else if (k == "adjoint_xcd") {
        if (value != 0 && value != 1) return MX_ERR_ARG;
        c->adj_xcd = (int)value;
    } else if (k == "adjoint_trace_rolled") {
        if (value != 0 && value != 1) return MX_ERR_ARG;
        c->adj_trace_rolled = (int)value;
    } else if (k == "adjoint_pow2_index") {
        if (value != 0 && value != 1) return MX_ERR_ARG;
        c->adj_pow2 = (int)value;
    } else {
        return MX_ERR_ARG;
    }
    return MX_OK;
}

int mx_adv_prof_read(mx_adv_ctx* c, int idx, const char** name, double* total_ms,
                     long long* launches) {
    if (!c) return MX_ERR_ARG;
    if (!c->prof_recs.empty()) {
        MX_HIP_CHECK(hipStreamSynchronize(c->stream));
        for (const auto& r : c->prof_recs) {
            float ms = 0.f;
            MX_HIP_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
            size_t k = 0;
            while (k < c->prof_names.size() && c->prof_names[k] != r.name) ++k;
            if (k == c->prof_names.size()) {
                c->prof_names.push_back(r.name);
                c->prof_ms.push_back(0.0);
                c->prof_n.push_back(0);
            }
            c->prof_ms[k] += ms;
            c->prof_n[k] += 1;
        }
        c->prof_recs.clear();
        c->ev_used = 0;
    }
    if (idx < 0) return (int)c->prof_names.size();
    if (idx >= (int)c->prof_names.size() || !name || !total_ms || !launches) return MX_ERR_ARG;
    *name = c->prof_names[idx].c_str();
    *total_ms = c->prof_ms[idx];
    *launches = c->prof_n[idx];
    return MX_OK;
}

int mx_adv_destroy(mx_adv_ctx* c) {
    if (!c) return MX_OK;
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    c->ev_pool.clear();
    (void)hipStreamSynchronize(c->stream);
    for (int lev = 0; lev < MX_MAX_LEV; ++lev) flux_reg_free(c, lev);
    if (c->ws) (void)hipFree(c->ws);
    if (c->adj_buf) (void)hipFree(c->adj_buf);
    for (auto& t : c->tab_slots) {
        free_tab_slot(t);
        if (t.done) (void)hipEventDestroy(t.done);
    }
    free_tab_cache(c);
    delete c;
    return MX_OK;
}

int mx_adv_set_level(mx_adv_ctx* c, int lev, int nbox, const mx_box* boxes) {
    if (!c || lev < 0 || lev > c->p.finest_level || nbox <= 0 || !boxes) return MX_ERR_ARG;
    if (nbox > 65535) return MX_ERR_ARG;
    c->grids[lev].clear();
    for (int b = 0; b < nbox; ++b) {
        mx_box bx = boxes[b];
        for (int d = 0; d < 3; ++d) {
            if (d >= c->p.dim) {
                bx.lo[d] = bx.hi[d] = 0;
            } else if (bx.lo[d] > bx.hi[d] || bx.lo[d] < c->p.domain_lo[lev][d] ||
                       bx.hi[d] > c->p.domain_hi[lev][d]) {
                c->grids[lev].clear();
                return MX_ERR_BOX;
            }
        }
        c->grids[lev].push_back(bx);
    }
    // a BoxArray's boxes are disjoint
    for (int a = 0; a < nbox; ++a)
        for (int b = a + 1; b < nbox; ++b) {
            const mx_box &x = c->grids[lev][a], &y = c->grids[lev][b];
            bool overlap = true;
            for (int d = 0; d < 3; ++d)
                if (std::max(x.lo[d], y.lo[d]) > std::min(x.hi[d], y.hi[d])) overlap = false;
            if (overlap) {
                c->grids[lev].clear();
                return MX_ERR_BOX;
            }
        }
    // plans depend on the BoxArray
    for (auto it = c->plans.begin(); it != c->plans.end();) {
        if (std::get<0>(it->first) == lev)
            it = c->plans.erase(it);
        else
            ++it;
    }
    for (auto it = c->fill_plans.begin(); it != c->fill_plans.end();) {
        if (std::get<0>(it->first) == lev)
            it = c->fill_plans.erase(it);
        else
            ++it;
    }
    // reflux_type 2: the level's flux register, Nscal components, as the
    // reference makes it with every new level (MaestroInit.cpp:389-392)
    if (lev > 0 && c->p.reflux_type == 2) return flux_reg_define(c, lev, c->p.nscal);
    return MX_OK;
}

int mx_adv_set_dt(mx_adv_ctx* c, double dt, double rel_eps) {
    if (!c) return MX_ERR_ARG;
    c->p.dt = dt;
    c->p.rel_eps = rel_eps;
    return MX_OK;
}

int mx_adv_set_stream(mx_adv_ctx* c, void* s) {
    if (!c) return MX_ERR_ARG;
    c->stream = (hipStream_t)s;
    return MX_OK;
}

int mx_adv_sync(mx_adv_ctx* c) {
    if (!c) return MX_ERR_ARG;
    MX_HIP_CHECK(hipStreamSynchronize(c->stream));
    MX_HIP_CHECK(hipGetLastError());
    return MX_OK;
}

const char* mx_adv_error_string(int code) {
    switch (code) {
        case MX_OK: return "ok";
        case MX_ERR_ARG: return "bad argument";
        case MX_ERR_BOX: return "FAB does not cover the region the stage needs";
        case MX_ERR_BC: return "invalid boundary type";
        case MX_ERR_HIP: return "HIP runtime failure";
        case MX_ERR_UNSUPPORTED: return "unsupported configuration";
        case MX_ERR_NOMEM: return "workspace allocation failed";
        case MX_ERR_NOLIB: return "no usable GPU";
        case MX_ERR_IO: return "file could not be created or written";
        case MX_ERR_SCRATCH:
            return "the call needs more device memory than the no-alloc scratch (256 MiB) holds";
        default: return "unknown error";
    }
}

int mx_adv_fill_face_ghosts(mx_adv_ctx* c, int lev, mx_face_multifab* uedge) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    rc = check_face(c, lev, uedge, 1, 1);
    if (rc) return rc;
    return fill_face_ghosts(c, lev, uedge);
}

int mx_adv_addw0(mx_adv_ctx* c, int lev, mx_face_multifab* uedge, const mx_face_multifab* w0mac,
                 const mx_basestate* w0, double mult) {
    int rc = check_level(c, lev);
    if (rc) return rc;
    rc = check_face(c, lev, uedge, 1, 1);
    if (rc) return rc;
    if (c->p.spherical) {
        rc = check_face(c, lev, w0mac, 1, 0);
        if (rc) return rc;
    } else if (!w0 || !w0->p || w0->len < c->p.domain_hi[lev][c->p.dim - 1] + 2) {
        return MX_ERR_ARG;
    }
    return addw0(c, lev, uedge, w0mac, w0, mult);
}

int mx_adv_cell_to_edge(mx_adv_ctx* c, int lev, int nr, const mx_basestate* s0_cell,
                        mx_basestate* s0_edge) {
    if (!c || !s0_cell || !s0_edge || !s0_cell->p || !s0_edge->p || nr < 4) return MX_ERR_ARG;
    if (s0_cell->len < nr || s0_edge->len < nr + 1) return MX_ERR_ARG;
    const int block = 256;
    MX_LAUNCH(c, "k_cell_to_edge", k_cell_to_edge,  dim3((nr + 1 + block - 1) / block), dim3(block), s0_cell->p + (long long)lev * s0_cell->len,
                       s0_edge->p + (long long)lev * s0_edge->len, nr);

    return MX_OK;
}

int mx_adv_bc_setup(const mx_adv_ctx* c, mx_bcrec* bcs_s, mx_bcrec* bcs_u);

int mx_adv_ppm(mx_adv_ctx* c, int lev, const mx_box* bx, const mx_fab* s, const mx_fab* u,
               const mx_fab* v, const mx_fab* w, mx_fab* Ip, mx_fab* Im, const mx_bcrec* bcs,
               int nbcs, int is_umac, int comp, int bccomp) {
    if (!c || !bx || !s || !u || !v || !Ip || !Im || !bcs) return MX_ERR_ARG;
    if (lev < 0 || lev > c->p.finest_level) return MX_ERR_ARG;
    const int dim = c->p.dim;
    if (dim == 3 && !w) return MX_ERR_ARG;
    if (bccomp < 0 || bccomp >= nbcs || comp < 0) return MX_ERR_ARG;
    if (c->p.ppm_type != 1 && c->p.ppm_type != 2) return MX_ERR_UNSUPPORTED;
    HBox r = hbox(*bx);
    if (dim == 2) r.lo[2] = r.hi[2] = 0;
    const int g = c->p.ppm_type == 2 ? 3 : 2;
    HBox sr = grow(r, dim, g);
    if (!covers(*s, sr, comp + 1) || !covers(*Ip, r, dim) || !covers(*Im, r, dim))
        return MX_ERR_BOX;
    for (int d = 0; d < dim; ++d) {
        const mx_fab* vf = d == 0 ? u : (d == 1 ? v : w);
        HBox vr = is_umac ? nodal(r, d) : r;
        if (!covers(*vf, vr, 1)) return MX_ERR_BOX;
    }
    PpmUnitArgs a;
    a.s = to_da(*s);
    a.u = to_da(*u);
    a.v = to_da(*v);
    a.w = w ? to_da(*w) : null_da();
    a.Ip = to_da(*Ip);
    a.Im = to_da(*Im);
    for (int d = 0; d < 3; ++d) {
        a.lo[d] = r.lo[d];
        a.hi[d] = r.hi[d];
        a.bclo[d] = bcs[bccomp].lo[d];
        a.bchi[d] = bcs[bccomp].hi[d];
    }
    a.comp = comp;
    a.is_umac = is_umac;
    KGeom kg = kgeom(c, lev);
    const int block = 256;
    dim3 gr = grid_for(npts(r), 1, dim, block);
    if (c->p.ppm_type == 1)
        MX_LAUNCH(c, "k_ppm_unit", k_ppm_unit<1>,  gr, dim3(block), a, kg);
    else
        MX_LAUNCH(c, "k_ppm_unit", k_ppm_unit<2>,  gr, dim3(block), a, kg);

    return MX_OK;
}

int mx_adv_slope(mx_adv_ctx* c, int lev, int dir, const mx_box* bx, const mx_fab* s, mx_fab* sl,
                 const mx_bcrec* bcs, int nbcs, int ncomp, int bc_start_comp, int slope_order) {
    if (!c || !bx || !s || !sl || !bcs) return MX_ERR_ARG;
    if (lev < 0 || lev > c->p.finest_level || dir < 0 || dir >= c->p.dim) return MX_ERR_ARG;
    if (ncomp <= 0 || ncomp > MX_MAX_SCAL || bc_start_comp < 0 || bc_start_comp + ncomp > nbcs)
        return MX_ERR_ARG;
    HBox r = hbox(*bx);
    if (c->p.dim == 2) r.lo[2] = r.hi[2] = 0;
    HBox sr = grow_dir(r, dir, 2);
    if (!covers(*s, sr, ncomp) || !covers(*sl, r, ncomp)) return MX_ERR_BOX;
    SlopeUnitArgs a;
    a.s = to_da(*s);
    a.sl = to_da(*sl);
    for (int d = 0; d < 3; ++d) {
        a.lo[d] = r.lo[d];
        a.hi[d] = r.hi[d];
    }
    for (int n = 0; n < ncomp; ++n) {
        a.bclo[n] = bcs[bc_start_comp + n].lo[dir];
        a.bchi[n] = bcs[bc_start_comp + n].hi[dir];
    }
    a.order = slope_order;
    a.ilo = c->p.domain_lo[lev][dir];
    a.ihi = c->p.domain_hi[lev][dir];
    const int block = 256;
    dim3 gr = grid_for(npts(r), 1, ncomp, block);
    if (dir == 0)
        MX_LAUNCH(c, "k_slope_unit", k_slope_unit<0>,  gr, dim3(block), a);
    else if (dir == 1)
        MX_LAUNCH(c, "k_slope_unit", k_slope_unit<1>,  gr, dim3(block), a);
    else
        MX_LAUNCH(c, "k_slope_unit", k_slope_unit<2>,  gr, dim3(block), a);

    return MX_OK;
}

}  // extern "C"
