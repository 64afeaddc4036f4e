// mx_noalloc.hip -- the advection pass on one level with no device
// allocation: mx_adv_advance_pass_noalloc.
//
// mx_adv_advance_pass keeps its box tables in device copies it allocates
// (cached per context) and VelPred's staged temporaries in a workspace it
// grows on demand.  This entry point runs the same stages and kernels with
// both placed in `scratch`, a fixed device array of this library
// (kScratchBytes per GPU, held until the process ends):
//
//   [0, kTabBytes)              the call's box tables, one region each;
//   [kTabBytes, kScratchBytes)  the workspace (VelPred's Ip/Im and
//                               transverse states).
//
// A dry run of the pass's host logic (no launch, no copy) sizes both before
// anything is queued, so a call that does not fit returns MX_ERR_SCRATCH with
// nothing launched.  Tables reach the device by stream-ordered copies from
// host memory the context keeps until the next such call, which waits for
// the stream first.  Every call on one GPU shares that GPU's array and
// overwrites it: calls that overlap on one GPU (other streams, other
// threads) race.
#include <mutex>

#include "mx_internal.H"

using namespace mxh;

constexpr size_t kScratchBytes = (size_t)256 << 20;
constexpr size_t kTabBytes = (size_t)16 << 20;
constexpr size_t kWsBytes = kScratchBytes - kTabBytes;

// one copy per GPU; host code uses the address hipGetSymbolAddress gives
// for the current device, never this name
static __device__ __align__(256) char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
char* g_scratch[kMaxDevices] = {};

// the device address of `scratch` on device `dev` (the current device),
// looked up once per device
int scratch_address(int dev, char** out) {
    if (dev < 0 || dev >= kMaxDevices) return MX_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MX_ERR_HIP;
        g_scratch[dev] = static_cast<char*>(p);
    }
    *out = g_scratch[dev];
    return MX_OK;
}

// restores the context's normal placement whatever path leaves the call
struct ExtGuard {
    mx_adv_ctx* c;
    int prof_on;
    explicit ExtGuard(mx_adv_ctx* cc) : c(cc), prof_on(cc->prof_on) { c->prof_on = 0; }
    ~ExtGuard() {
        c->dry = 0;
        c->ext_tab = nullptr;
        c->ext_tab_cap = c->ext_tab_off = 0;
        c->ext_ws = nullptr;
        c->ext_ws_cap = 0;
        c->prof_on = prof_on;
    }
};

}  // namespace

extern "C" {

int mx_adv_advance_pass_noalloc(mx_adv_ctx* c, int lev, const mx_adv_pass_args* a) {
    if (!c || !a) return MX_ERR_ARG;
    if (c->tab_mode != 0) return MX_ERR_ARG;  // not inside a graph recording
    int dev = -1;
    MX_HIP_CHECK(hipGetDevice(&dev));
    if (dev != c->device) return MX_ERR_ARG;
    char* base = nullptr;
    int rc = scratch_address(dev, &base);
    if (rc) return rc;
    // the host copies of the previous call's tables may still be in flight
    if (!c->ext_host.empty()) {
        MX_HIP_CHECK(hipStreamSynchronize(c->stream));
        c->ext_host.clear();
    }
    ExtGuard guard(c);
    c->dry = 1;
    c->dry_tab = c->dry_ws = 0;
    rc = advance_pass_levels(c, lev, 1, a);
    c->dry = 0;
    if (rc) return rc;
    if (c->dry_tab > kTabBytes || c->dry_ws > kWsBytes) return MX_ERR_SCRATCH;
    c->ext_tab = base;
    c->ext_tab_cap = kTabBytes;
    c->ext_tab_off = 0;
    c->ext_ws = base + kTabBytes;
    c->ext_ws_cap = kWsBytes;
    return advance_pass_levels(c, lev, 1, a);
}

}  // extern "C"
