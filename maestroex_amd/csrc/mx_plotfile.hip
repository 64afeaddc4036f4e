// mx_plotfile.hip -- amrex::WriteMultiLevelPlotfile as Maestro::WritePlotFile
// calls it (MaestroPlot.cpp:99-116): the AMReX plotfile directory of one
// multi-level cell-centred field (format "HyperCLaw-V1.1").
//
//   <dir>/Header                  the hierarchy: variables, geometry, boxes
//   <dir>/Level_<l>/Cell_H        VisMF header (version 1): BoxArray, the
//                                 FAB offsets in Cell_D_00000, per-FAB
//                                 per-component minima and maxima
//   <dir>/Level_<l>/Cell_D_00000  every FAB of the level: a text header
//                                 "FAB <real descriptor><box> <ncomp>\n"
//                                 and the valid cells as little-endian
//                                 IEEE doubles, i fastest, component outermost
//
// Host code: each FAB is copied to a pinned host buffer on the context
// stream (the copy waits for the kernels before it); the valid cells are
// gathered and written there.
#include <sys/stat.h>
#include <sys/types.h>

#include <cerrno>
#include <cstdio>
#include <string>

#include "mx_internal.H"

using namespace mxh;

namespace {

std::string box_str(const int lo[3], const int hi[3], int dim) {
    std::string s = "((";
    for (int d = 0; d < dim; ++d) s += std::to_string(lo[d]) + (d + 1 < dim ? "," : "");
    s += ") (";
    for (int d = 0; d < dim; ++d) s += std::to_string(hi[d]) + (d + 1 < dim ? "," : "");
    s += ") (";
    for (int d = 0; d < dim; ++d) s += std::string("0") + (d + 1 < dim ? "," : "");
    s += "))";
    return s;
}

// a double as C++ streams print it with precision(17) and no floatfield
std::string g17(double v) {
    char b[64];
    std::snprintf(b, sizeof(b), "%.17g", v);
    return b;
}

// as VisMF writes minima and maxima: scientific, precision 16
std::string e16(double v) {
    char b[64];
    std::snprintf(b, sizeof(b), "%.16e", v);
    return b;
}

// the refinement ratio between levels l and l + 1: one integer in every
// direction, or 0
int ref_ratio(const mx_adv_ctx* c, int l) {
    int r = 0;
    for (int d = 0; d < c->p.dim; ++d) {
        const int nc = c->p.domain_hi[l][d] - c->p.domain_lo[l][d] + 1;
        const int nf = c->p.domain_hi[l + 1][d] - c->p.domain_lo[l + 1][d] + 1;
        if (nc < 1 || nf % nc != 0 || (d > 0 && nf / nc != r)) return 0;
        r = nf / nc;
    }
    return r;
}

int make_dir(const std::string& p) {
    if (mkdir(p.c_str(), 0755) != 0 && errno != EEXIST) return MX_ERR_IO;
    return MX_OK;
}

}  // namespace

extern "C" {

int mx_adv_write_plotfile(mx_adv_ctx* c, const char* dir, int nlev, const mx_multifab* const* mf,
                          int ncomp, const char* const* varnames, double time,
                          const int* level_steps, const double* prob_lo, const double* prob_hi) {
    if (!c || !dir || !mf || !varnames || !level_steps || !prob_lo || !prob_hi) return MX_ERR_ARG;
    if (nlev < 1 || nlev - 1 > c->p.finest_level || ncomp < 1) return MX_ERR_ARG;
    const int dim = c->p.dim;
    for (int l = 0; l < nlev; ++l) {
        const auto& g = c->grids[l];
        if (!mf[l] || mf[l]->nfab != (int)g.size() || g.empty()) return MX_ERR_ARG;
        for (size_t b = 0; b < g.size(); ++b) {
            const mx_fab& fab = mf[l]->fab[b];
            if (!covers(fab, hbox(g[b]), ncomp)) return MX_ERR_BOX;
            // the copies take each component's planes whole: a dense FAB
            const long long nyf = fab.end[1] - fab.begin[1], nzf = fab.end[2] - fab.begin[2];
            if (fab.jstride < fab.end[0] - fab.begin[0] || fab.kstride < fab.jstride * nyf ||
                (ncomp > 1 && fab.nstride < fab.kstride * nzf))
                return MX_ERR_BOX;
        }
        if (l > 0 && ref_ratio(c, l - 1) < 1) return MX_ERR_ARG;
    }
    for (int n = 0; n < ncomp; ++n)
        if (!varnames[n]) return MX_ERR_ARG;
    const std::string root(dir);
    int rc = make_dir(root);
    if (rc) return rc;

    // Header
    {
        std::string h = "HyperCLaw-V1.1\n" + std::to_string(ncomp) + "\n";
        for (int n = 0; n < ncomp; ++n) h += std::string(varnames[n]) + "\n";
        h += std::to_string(dim) + "\n" + g17(time) + "\n" + std::to_string(nlev - 1) + "\n";
        for (int d = 0; d < dim; ++d) h += g17(prob_lo[d]) + " ";
        h += "\n";
        for (int d = 0; d < dim; ++d) h += g17(prob_hi[d]) + " ";
        h += "\n";
        for (int l = 0; l + 1 < nlev; ++l) h += std::to_string(ref_ratio(c, l)) + " ";
        h += "\n";
        for (int l = 0; l < nlev; ++l)
            h += box_str(c->p.domain_lo[l], c->p.domain_hi[l], dim) + " ";
        h += "\n";
        for (int l = 0; l < nlev; ++l) h += std::to_string(level_steps[l]) + " ";
        h += "\n";
        for (int l = 0; l < nlev; ++l) {
            for (int d = 0; d < dim; ++d) h += g17(c->p.dx[l][d]) + " ";
            h += "\n";
        }
        h += "0\n0\n";  // Cartesian coordinates; boundary width 0
        for (int l = 0; l < nlev; ++l) {
            const auto& g = c->grids[l];
            h += std::to_string(l) + " " + std::to_string(g.size()) + " " + g17(time) + "\n";
            h += std::to_string(level_steps[l]) + "\n";
            for (const mx_box& b : g)
                for (int d = 0; d < dim; ++d) {
                    const double dx = c->p.dx[l][d];
                    const int lo = b.lo[d] - c->p.domain_lo[l][d];
                    const int hi = b.hi[d] - c->p.domain_lo[l][d];
                    h += g17(prob_lo[d] + dx * lo) + " " + g17(prob_lo[d] + dx * (hi + 1)) + "\n";
                }
            h += "Level_" + std::to_string(l) + "/Cell\n";
        }
        FILE* f = std::fopen((root + "/Header").c_str(), "wb");
        if (!f) return MX_ERR_IO;
        const bool ok = std::fwrite(h.data(), 1, h.size(), f) == h.size();
        if (std::fclose(f) != 0 || !ok) return MX_ERR_IO;
    }

    // levels: each FAB's components are copied whole (ghosts included, one
    // contiguous copy per component) into a pinned staging buffer, and the
    // valid cells are gathered from there on the host
    size_t stage_cap = 0;
    for (int l = 0; l < nlev; ++l)
        for (int b = 0; b < mf[l]->nfab; ++b) {
            const mx_fab& fab = mf[l]->fab[b];
            const size_t comp = (size_t)fab.kstride * (size_t)(fab.end[2] - fab.begin[2]);
            stage_cap = std::max(stage_cap, comp * (size_t)ncomp);
        }
    double* stage = nullptr;
    if (hipHostMalloc((void**)&stage, stage_cap * sizeof(double), hipHostMallocDefault) !=
        hipSuccess)
        return MX_ERR_NOMEM;
    struct StageFree {
        double* p;
        ~StageFree() { (void)hipHostFree(p); }
    } stage_free{stage};
    std::vector<double> host;
    for (int l = 0; l < nlev; ++l) {
        const std::string ldir = root + "/Level_" + std::to_string(l);
        if ((rc = make_dir(ldir))) return rc;
        const auto& g = c->grids[l];
        FILE* f = std::fopen((ldir + "/Cell_D_00000").c_str(), "wb");
        if (!f) return MX_ERR_IO;
        std::vector<long long> offs;
        std::vector<std::vector<double>> mins, maxs;
        long long off = 0;
        bool ok = true, hip_ok = true;
        for (size_t b = 0; b < g.size() && ok; ++b) {
            const mx_fab& fab = mf[l]->fab[b];
            const mx_box& bx = g[b];
            const long long nx = bx.hi[0] - bx.lo[0] + 1;
            const long long ny = dim > 1 ? bx.hi[1] - bx.lo[1] + 1 : 1;
            const long long nz = dim > 2 ? bx.hi[2] - bx.lo[2] + 1 : 1;
            const size_t comp = (size_t)fab.kstride * (size_t)(fab.end[2] - fab.begin[2]);
            for (int n = 0; n < ncomp && hip_ok; ++n)
                hip_ok = hipMemcpyAsync(stage + (size_t)n * comp, fab.p + (size_t)n * fab.nstride,
                                        comp * sizeof(double), hipMemcpyDeviceToHost,
                                        c->stream) == hipSuccess;
            if (!hip_ok || hipStreamSynchronize(c->stream) != hipSuccess) {
                hip_ok = ok = false;
                break;
            }
            host.resize((size_t)(nx * ny * nz * ncomp));
            const size_t npt = (size_t)(nx * ny * nz);
            for (int n = 0; n < ncomp; ++n)
                for (long long k = 0; k < nz; ++k)
                    for (long long j = 0; j < ny; ++j) {
                        const double* src = stage + (size_t)n * comp +
                                            (bx.lo[0] - fab.begin[0]) +
                                            (bx.lo[1] + j - fab.begin[1]) * fab.jstride +
                                            (bx.lo[2] + k - fab.begin[2]) * fab.kstride;
                        std::memcpy(host.data() + (size_t)n * npt + (size_t)((k * ny + j) * nx),
                                    src, (size_t)nx * sizeof(double));
                    }
            std::string head = "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (1 2 3 4 5 6 7 8)))";
            int lo[3] = {bx.lo[0], bx.lo[1], bx.lo[2]}, hi[3] = {bx.hi[0], bx.hi[1], bx.hi[2]};
            head += box_str(lo, hi, dim) + " " + std::to_string(ncomp) + "\n";
            offs.push_back(off);
            ok = std::fwrite(head.data(), 1, head.size(), f) == head.size() &&
                 std::fwrite(host.data(), sizeof(double), host.size(), f) == host.size();
            off += (long long)head.size() + (long long)(host.size() * sizeof(double));
            std::vector<double> mn(ncomp), mx(ncomp);
            for (int n = 0; n < ncomp; ++n) {
                double a = host[(size_t)n * npt], z = a;
                for (size_t q = 1; q < npt; ++q) {
                    const double v = host[(size_t)n * npt + q];
                    a = v < a ? v : a;
                    z = v > z ? v : z;
                }
                mn[n] = a;
                mx[n] = z;
            }
            mins.push_back(mn);
            maxs.push_back(mx);
        }
        if (std::fclose(f) != 0 || !ok) return hip_ok ? MX_ERR_IO : MX_ERR_HIP;

        std::string h = "1\n1\n" + std::to_string(ncomp) + "\n0\n";
        h += "(" + std::to_string(g.size()) + " 0\n";
        for (const mx_box& b : g) {
            int lo[3] = {b.lo[0], b.lo[1], b.lo[2]}, hi[3] = {b.hi[0], b.hi[1], b.hi[2]};
            h += box_str(lo, hi, dim) + "\n";
        }
        h += ")\n" + std::to_string(g.size()) + "\n";
        for (long long o : offs) h += "FabOnDisk: Cell_D_00000 " + std::to_string(o) + "\n";
        h += "\n";
        for (int which = 0; which < 2; ++which) {
            const auto& v = which ? maxs : mins;
            h += std::to_string(g.size()) + "," + std::to_string(ncomp) + "\n";
            for (const auto& row : v) {
                for (double x : row) h += e16(x) + ",";
                h += "\n";
            }
            h += "\n";
        }
        FILE* fh = std::fopen((ldir + "/Cell_H").c_str(), "wb");
        if (!fh) return MX_ERR_IO;
        const bool okh = std::fwrite(h.data(), 1, h.size(), fh) == h.size();
        if (std::fclose(fh) != 0 || !okh) return MX_ERR_IO;
    }
    return MX_OK;
}

}  // extern "C"
