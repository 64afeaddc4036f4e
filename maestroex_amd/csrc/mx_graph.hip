// mx_graph.hip -- the pass over a whole level hierarchy: eager, captured as
// one HIP graph, and the autotuner of the fused edge-state kernel's launch
// shape.
//
// Graph capture cannot contain the per-call box-table uploads (host buffers
// that the next call overwrites) nor any workspace allocation.  Creation
// therefore runs the pass three times: eagerly to size the workspace,
// eagerly again with every table recorded, then under stream capture with
// each table taken from a device arena that holds the recorded copies.  The graph owns
// that arena and the workspace its kernels use.
#include <algorithm>
#include <cstring>

#include "mx_internal.H"

using namespace mxh;

struct mx_adv_graph {
    mx_adv_ctx* c = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    char* arena = nullptr;
    char* ws = nullptr;
    size_t ws_cap = 0;
    // what a re-capture needs: the per-step values the kernels take by value
    // at capture time, the pass arguments with graph-owned copies of the
    // caller's FAB descriptor arrays, and the recorded box tables
    double dt = 0.0, rel_eps = 0.0;
    int nlev = 0;
    std::vector<mx_adv_pass_args> args;
    std::vector<std::vector<mx_fab>> fabs;
    std::vector<std::vector<char>> tabs;
    std::vector<size_t> offs;
};

namespace {

void graph_free(mx_adv_graph* g) {
    if (!g) return;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->arena) (void)hipFree(g->arena);
    if (g->ws) (void)hipFree(g->ws);
    delete g;
}

struct TabModeGuard {
    mx_adv_ctx* c;
    int prof;
    explicit TabModeGuard(mx_adv_ctx* ctx) : c(ctx), prof(ctx->prof_on) { c->prof_on = 0; }
    ~TabModeGuard() {
        c->tab_mode = 0;
        c->rec_tabs.clear();
        c->rec_offs.clear();
        c->rec_arena = nullptr;
        c->rec_next = 0;
        c->prof_on = prof;
    }
};

// graph-owned copies of the FAB descriptor arrays the pass arguments point
// at, so that a later re-capture does not read caller memory that may be gone
void keep_mf(mx_adv_graph* g, mx_multifab& m) {
    if (m.nfab <= 0 || !m.fab) return;
    g->fabs.emplace_back(m.fab, m.fab + m.nfab);
    m.fab = g->fabs.back().data();
}
void keep_face(mx_adv_graph* g, mx_face_multifab& f) {
    for (int d = 0; d < 3; ++d) keep_mf(g, f.dir[d]);
}
void keep_args(mx_adv_graph* g, const mx_adv_pass_args* args, int nlev) {
    g->args.assign(args, args + nlev);
    size_t n = 0;
    for (int i = 0; i < nlev; ++i) n += 6 + 1 + 8 * 3;
    g->fabs.reserve(n);  // no reallocation: the data() pointers stay valid
    for (mx_adv_pass_args& a : g->args) {
        for (mx_multifab* m : {&a.utilde, &a.ufull, &a.vel_force, &a.w0_cart, &a.scal,
                               &a.scal_force, &a.etarhoflux})
            keep_mf(g, *m);
        for (mx_face_multifab* f :
             {&a.w0mac, &a.r0mac, &a.rh0mac, &a.h0mac, &a.utrans, &a.umac, &a.sedge, &a.sflux})
            keep_face(g, *f);
    }
}

// Stream capture of the pass with every box table taken from the graph's
// arena and the graph's workspace (so the captured addresses are those of
// the first capture); on a stream of our own when the context uses the null
// stream.  The caller holds a TabModeGuard.
int capture(mx_adv_graph* g, hipGraph_t* out) {
    mx_adv_ctx* c = g->c;
    *out = nullptr;
    c->rec_tabs = g->tabs;
    c->rec_offs = g->offs;
    c->rec_arena = g->arena;
    c->rec_next = 0;
    char* ws_saved = c->ws;
    const size_t cap_saved = c->ws_cap, off_saved = c->ws_off;
    c->ws = g->ws;
    c->ws_cap = g->ws_cap;
    c->ws_off = 0;
    hipStream_t saved = c->stream;
    hipStream_t cap = saved;
    int rc = MX_OK;
    if (!cap && hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess) rc = MX_ERR_HIP;
    if (!rc) {
        c->stream = cap;
        c->tab_mode = 2;
        hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            rc = advance_pass_levels(c, 0, g->nlev, g->args.data());
            hipGraph_t graph = nullptr;
            e = hipStreamEndCapture(cap, &graph);
            *out = graph;
        }
        if (e != hipSuccess) rc = rc ? rc : MX_ERR_HIP;
        if (!rc && c->rec_next != c->rec_tabs.size()) rc = MX_ERR_ARG;
        c->stream = saved;
        if (!saved) (void)hipStreamDestroy(cap);
    }
    // the workspace must not have moved: the graph owns it
    if (!rc && c->ws != g->ws) rc = MX_ERR_ARG;
    c->ws = ws_saved;
    c->ws_cap = cap_saved;
    c->ws_off = off_saved;
    return rc;
}

}  // namespace

extern "C" {

int mx_adv_advance_pass_levels(mx_adv_ctx* c, int nlev, const mx_adv_pass_args* args) {
    if (!c || !args || nlev < 1) return MX_ERR_ARG;
    return advance_pass_levels(c, 0, nlev, args);
}

int mx_adv_pass_graph_create(mx_adv_ctx* c, int nlev, const mx_adv_pass_args* args,
                             mx_adv_graph** out) {
    if (!c || !args || !out || nlev < 1) return MX_ERR_ARG;
    *out = nullptr;
    int rc;
    mx_adv_graph* g = new mx_adv_graph();
    g->c = c;
    g->nlev = nlev;
    g->dt = c->p.dt;
    g->rel_eps = c->p.rel_eps;
    keep_args(g, args, nlev);
    TabModeGuard guard(c);

    // 1) an eager run sizes the workspace (its address then stays fixed), a
    //    second one records every box table
    if ((rc = advance_pass_levels(c, 0, nlev, g->args.data()))) {
        graph_free(g);
        return rc;
    }
    c->tab_mode = 1;
    c->rec_tabs.clear();
    if ((rc = advance_pass_levels(c, 0, nlev, g->args.data()))) {
        graph_free(g);
        return rc;
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        graph_free(g);
        return MX_ERR_HIP;
    }
    c->tab_mode = 0;

    // 2) the recorded tables in one device arena
    size_t total = 0;
    for (const auto& t : c->rec_tabs) {
        g->offs.push_back(total);
        total += (t.size() + 255) & ~(size_t)255;
    }
    g->tabs.swap(c->rec_tabs);
    if (total) {
        if (hipMalloc((void**)&g->arena, total) != hipSuccess) {
            g->arena = nullptr;
            graph_free(g);
            return MX_ERR_NOMEM;
        }
        std::vector<char> host(total, 0);
        for (size_t i = 0; i < g->tabs.size(); ++i)
            if (!g->tabs[i].empty())
                std::memcpy(host.data() + g->offs[i], g->tabs[i].data(), g->tabs[i].size());
        if (hipMemcpy(g->arena, host.data(), total, hipMemcpyHostToDevice) != hipSuccess) {
            graph_free(g);
            return MX_ERR_HIP;
        }
    }
    // the graph's kernels use the context workspace as it is now: hand it to
    // the graph so that later eager calls allocate their own
    g->ws = c->ws;
    g->ws_cap = c->ws_cap;
    c->ws = nullptr;
    c->ws_cap = 0;
    c->ws_off = 0;

    // 3) capture, then instantiate
    hipGraph_t graph = nullptr;
    rc = capture(g, &graph);
    g->graph = graph;
    if (!rc && hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
        rc = MX_ERR_HIP;
    if (rc) {
        graph_free(g);
        return rc;
    }
    *out = g;
    return MX_OK;
}

int mx_adv_pass_graph_launch(mx_adv_graph* g) {
    if (!g || !g->exec) return MX_ERR_ARG;
    mx_adv_ctx* c = g->c;
    if (c->p.dt != g->dt || c->p.rel_eps != g->rel_eps) {
        // dt and rel_eps changed since capture (mx_adv_set_dt, every step of
        // the reference's time loop): the kernels take them by value, so the
        // pass is captured again with the new values and the executable
        // graph updated in place (re-instantiated if the update is refused)
        hipGraph_t fresh = nullptr;
        int rc;
        {
            TabModeGuard guard(c);
            rc = capture(g, &fresh);
        }
        if (rc) {
            if (fresh) (void)hipGraphDestroy(fresh);
            return rc;
        }
        hipGraphNode_t err_node = nullptr;
        hipGraphExecUpdateResult res;
        if (hipGraphExecUpdate(g->exec, fresh, &err_node, &res) != hipSuccess ||
            res != hipGraphExecUpdateSuccess) {
            (void)hipGetLastError();
            hipGraphExec_t ex = nullptr;
            if (hipGraphInstantiate(&ex, fresh, nullptr, nullptr, 0) != hipSuccess) {
                (void)hipGraphDestroy(fresh);
                return MX_ERR_HIP;
            }
            (void)hipGraphExecDestroy(g->exec);
            g->exec = ex;
        }
        (void)hipGraphDestroy(g->graph);
        g->graph = fresh;
        g->dt = c->p.dt;
        g->rel_eps = c->p.rel_eps;
    }
    if (hipGraphLaunch(g->exec, c->stream) != hipSuccess) return MX_ERR_HIP;
    return MX_OK;
}

int mx_adv_pass_graph_destroy(mx_adv_graph* g) {
    if (!g) return MX_OK;
    (void)hipStreamSynchronize(g->c->stream);
    graph_free(g);
    return MX_OK;
}

int mx_adv_autotune(mx_adv_ctx* c, int nlev, const mx_adv_pass_args* args,
                    mx_adv_tune_report* rep) {
    if (!c || !args || nlev < 1 || !rep) return MX_ERR_ARG;
    std::memset(rep, 0, sizeof(*rep));
    rep->best = -1;
    int rc;
    // one pass for the MAC velocities the edge states are traced with
    if ((rc = advance_pass_levels(c, 0, nlev, args))) return rc;
    if (!fused_edge_scal_ok(c, 0)) return MX_ERR_UNSUPPORTED;
    const mx_adv_params& p = c->p;
    std::vector<EdgeLevel> levs;
    for (int i = 0; i < nlev; ++i)
        levs.push_back({i, &args[i].scal, &args[i].sedge, &args[i].umac, &args[i].scal_force});
    const mx_bcrec* bs = c->bcs_s.data();
    const int tiles[] = {4, 2, 3, 0};
    const int tzs[] = {16, 32, 64, 128};
    const int saved_tile = c->fused_tile, saved_tz = c->fused_tz, saved_prof = c->prof_on;
    c->prof_on = 0;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return MX_ERR_HIP;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return MX_ERR_HIP;
    }
    rc = MX_OK;
    double best = 1e300;
    for (int tile : tiles) {
        for (int tz : tzs) {
            if (rep->ncand >= MX_TUNE_MAX) break;
            c->fused_tile = tile;
            c->fused_tz = tz;
            float ms_min = 1e30f;
            for (int rep_i = 0; rep_i < 3 && !rc; ++rep_i) {
                if (hipEventRecord(e0, c->stream) != hipSuccess) rc = MX_ERR_HIP;
                if (!rc)
                    rc = make_edge_scal_fused_levels(c, levs, 0, bs, p.first_spec, p.first_spec,
                                                     p.num_spec, 0);
                if (!rc && hipEventRecord(e1, c->stream) != hipSuccess) rc = MX_ERR_HIP;
                if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = MX_ERR_HIP;
                float ms = 0.f;
                if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = MX_ERR_HIP;
                if (!rc && rep_i > 0) ms_min = std::min(ms_min, ms);  // first run warms up
            }
            if (rc) break;
            const int k = rep->ncand++;
            rep->tile[k] = tile;
            rep->tz[k] = tz;
            rep->ms[k] = ms_min;
            if (ms_min < best) {
                best = ms_min;
                rep->best = k;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->prof_on = saved_prof;
    if (rc || rep->best < 0) {
        c->fused_tile = saved_tile;
        c->fused_tz = saved_tz;
        return rc ? rc : MX_ERR_HIP;
    }
    c->fused_tile = rep->tile[rep->best];
    c->fused_tz = rep->tz[rep->best];
    return MX_OK;
}

}  // extern "C"
