// Entry points that allocate no device memory, for callers that cannot call
// hipMalloc: code under HIP graph capture, tight serving loops, harnesses
// that forbid allocation.  Their device scratch is a file-scope array of
// fixed size, one copy per GPU, whose address is looked up once per device
// when a context is created there (hipGetSymbolAddress; the array's name is
// only a host-side stand-in for it).  Every call on a GPU shares that GPU's
// array and leaves the last call's contents in it; calls that overlap on one
// GPU race.
#include "igr_ctx.hpp"

#include <mutex>
#include <string>

using namespace igr_dev;

namespace {

constexpr int MAX_DEVICES = 64;

// Two regions of IGR_NOALLOC_MAX_GCELLS floats (64 MiB in all): the per-gcell
// congestion view, then each prefix view in turn.
static __device__ __align__(256) float scratch[2 * (size_t)IGR_NOALLOC_MAX_GCELLS];

std::mutex g_mu;
float* g_ptr[MAX_DEVICES] = {};
hipError_t g_err[MAX_DEVICES] = {};
bool g_done[MAX_DEVICES] = {};

float* scratch_of(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (device < 0 || device >= MAX_DEVICES || !g_done[device])
        throw std::logic_error("no-allocation scratch not set up on this device");
    if (g_err[device] != hipSuccess || !g_ptr[device])
        throw std::runtime_error(std::string("no-allocation scratch unavailable: ") + hipGetErrorString(g_err[device]));
    return g_ptr[device];
}

} // namespace

void igr_dev::noalloc_init(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (device < 0 || device >= MAX_DEVICES || g_done[device]) return;
    void* p = nullptr;
    g_err[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_ptr[device] = g_err[device] == hipSuccess ? static_cast<float*>(p) : nullptr;
    g_done[device] = true;
}

int igr_congestion_view_noalloc(igr_ctx* c, float* xsum, float* ysum) {
    if (!c || !xsum || !ysum) return fail(IGR_EINVAL, "null context or output");
    const int64_t q = c->g.XYc;
    if (q > IGR_NOALLOC_MAX_GCELLS)
        return fail(IGR_EINVAL, "grid of " + std::to_string(q) + " gcells exceeds the no-allocation limit of " +
                                    std::to_string((long long)IGR_NOALLOC_MAX_GCELLS) + " (X*Y)");
    try {
        int cur = -1;
        HIP_OK(hipGetDevice(&cur));
        if (cur != c->device) HIP_OK(hipSetDevice(c->device));
        float* base = scratch_of(c->device);
        float* view = base;
        float* out = base + IGR_NOALLOC_MAX_GCELLS;
        // The copy of each prefix view is ordered on the stream before the
        // next launch writes the same region.
        launch_congestion_map(c, view);
        launch_view_scan(c, view, out, 0);
        HIP_OK(hipMemcpyAsync(xsum, out, sizeof(float) * q, hipMemcpyDeviceToHost, c->stream));
        launch_view_scan(c, view, out, 1);
        HIP_OK(hipMemcpyAsync(ysum, out, sizeof(float) * q, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        if (cur != c->device) HIP_OK(hipSetDevice(cur));
    } catch (const std::logic_error& e) {
        return fail(IGR_ESTATE, e.what());
    } catch (const std::exception& e) {
        return fail(IGR_EHIP, e.what());
    }
    return IGR_OK;
}
