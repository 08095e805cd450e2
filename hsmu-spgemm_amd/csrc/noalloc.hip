// noalloc.hip -- the workspace of hsmu_spgemm_noalloc: one file-scope device array.
//
// hsmu_spgemm takes its workspace (B's tiles, per-row analysis, bins, row directories,
// wide-row work items) from the context's caching pool, which calls hipMalloc the first
// time a size is needed.  Callers that may not allocate device memory at all (code under
// stream capture, serving loops, harnesses that forbid it) use hsmu_spgemm_noalloc, whose
// workspace is this array instead: NOALLOC_SCRATCH_BYTES of device memory that every GPU
// the library runs on holds from module load until the process ends.
//
// In host code `scratch` names a host-side stand-in, not the device array: its device
// address is looked up once per device with hipGetSymbolAddress (scratch_setup, called by
// hsmu_ctx_create with the device current) and only that address is ever passed to
// kernels or copies.  Every call on one device shares the array, so scratch_lock
// serialises the calls of this process on a device.
#include <hip/hip_runtime.h>

#include <mutex>

#include "hsmu_internal.h"

static __device__ __align__(256) unsigned char scratch[hsmu::NOALLOC_SCRATCH_BYTES];

namespace hsmu {

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_setup_mu;
void *g_addr[MAX_DEVICES] = {};
std::mutex g_call_mu[MAX_DEVICES];
} // namespace

hipError_t scratch_setup(int device, void **ptr)
{
    *ptr = nullptr;
    if (device < 0 || device >= MAX_DEVICES)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_setup_mu);
    if (!g_addr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess)
            return e;
        if (cur != device)
            return hipErrorInvalidDevice; // the lookup must run with `device` current
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        if (!p)
            return hipErrorInvalidSymbol;
        g_addr[device] = p;
    }
    *ptr = g_addr[device];
    return hipSuccess;
}

std::mutex &scratch_lock(int device) { return g_call_mu[device >= 0 && device < MAX_DEVICES ? device : 0]; }

} // namespace hsmu
