// qpc_noalloc.hip -- the allocation-free observables entry point
// (qpc_observe_noalloc): its scratch array and the per-device resolution of
// the array's address.
#include "qpc_ctx.h"

#include <mutex>

// Scratch of qpc_observe_noalloc: the per-channel partial sums of the
// observables (qpc::kObs = 8 doubles per channel), in a file-scope device array instead
// of memory the library allocates.  Each GPU holds its own copy for the life
// of the process; host code reaches it only through the address
// hipGetSymbolAddress returns on that GPU (resolved once per device, when
// the first context is created there), never through the array's name.
static constexpr int64_t kNoAllocChannels = 65536;
static __device__ __align__(256) double scratch[qpc::kObs * kNoAllocChannels];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// resolves this device's copy of `scratch` (the device must be current)
int resolve_scratch(qpc_ctx* c) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (c->device < 0 || c->device >= kMaxDevices) {
        c->set_error("device index beyond the scratch table");
        return QPC_E_HIP;
    }
    if (g_scratch[c->device]) return QPC_OK;
    void* p = nullptr;
    QPC_HIP_CHECK(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[c->device] = static_cast<double*>(p);
    return QPC_OK;
}

}  // namespace

namespace qpc_host {
int resolve_scratch(qpc_ctx* c) { return ::resolve_scratch(c); }
}  // namespace qpc_host

extern "C" {

int qpc_observe_noalloc(qpc_ctx* c, int flags, double* out) {
    if (!c || !out || flags < 1 || flags > QPC_REC_ALL || c->batch) return QPC_E_ARG;
    if (c->ML > kNoAllocChannels) {
        c->set_error("qpc_observe_noalloc: more than 65536 channels (its scratch holds 8 x 65536 partial sums)");
        return QPC_E_SHAPE;
    }
    if (!c->have_radial || !c->have_state) return QPC_E_STATE;
    if (((flags & (QPC_REC_ACCEL | QPC_REC_ACCEL_XY)) && !c->have_accel) || ((flags & QPC_REC_OUTER) && !c->have_outer)) {
        c->set_error("band of a requested observable not uploaded");
        return QPC_E_STATE;
    }
    double* red = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        red = c->device >= 0 && c->device < kMaxDevices ? g_scratch[c->device] : nullptr;
    }
    if (!red) {
        c->set_error("scratch address not resolved on this device");
        return QPC_E_STATE;
    }
    // every partial the sum reads is written first (k_observe writes all
    // channels of every requested observable), so nothing needs clearing
    const qpc::ObsBands b{c->S, c->rpow1, c->accel, c->souter, c->y10};
    qpc::launch_observe(c->stream, b, flags, c->coef, red, out, c->M, c->L, c->m0, c->N, c->p);
    return launch_check(c, "observe_noalloc");
}

}  // extern "C"
