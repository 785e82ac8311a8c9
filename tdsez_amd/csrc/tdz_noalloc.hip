// Allocation-free entry point (no counterpart in the reference): the observables of
// tdz_observables for callers that cannot allocate device memory (code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc). The reductions' partial sums
// live in a file-scope device array instead of memory the call would allocate.
//
// Each GPU holds its own copy of the array; its device address is resolved once per device
// (noalloc_init, called by tdz_create with the context's device current) and only that
// address is handed to kernels and copies. Every call on a GPU shares that GPU's array:
// calls that overlap on one GPU (two host threads, two contexts on different streams)
// race and must be serialised by the caller. Nothing in the array needs to start at zero
// (every partial is written before it is read).
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "../../include/tdz.h"
#include "tdz_internal.hpp"

namespace tdz {

int observables_device(tdz_ctx *c, const double *psi, double out[6], dcomplex *scr);
int ctx_device(const tdz_ctx *c);
int fail(const std::string &msg);

static __device__ __align__(256) dcomplex scratch[noalloc_scratch_len()];   // 65.7 KB per GPU
constexpr int kMaxDevices = 64;
static dcomplex *g_scratch[kMaxDevices] = {};   // device address of scratch, per device

void noalloc_init()
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        throw std::runtime_error("noalloc_init: no current device (or more than 64)");
    if (g_scratch[dev]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch[dev] = static_cast<dcomplex *>(p);
}

}  // namespace tdz

extern "C" int tdz_observables_noalloc(tdz_ctx *ctx, const double *psi, double out[6])
{
    if (!ctx) return tdz::fail("null context");
    const int dev = tdz::ctx_device(ctx);
    if (dev < 0 || dev >= tdz::kMaxDevices || !tdz::g_scratch[dev])
        return tdz::fail("tdz_observables_noalloc: the static scratch of the context's device is not set up");
    return tdz::observables_device(ctx, psi, out, tdz::g_scratch[dev]);
}
