// stencil_dispatch_noalloc: stencil_dispatch for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  No reference counterpart.
//
// The dispatch's own device scratch lives in the file-scope array `scratch`
// below instead of the library's allocated workspace: the row form's
// shell-comparison flag (slot 2, 4 KiB part) and the streaming form's flags
// and x-shell table of both buffers (slot 3: 1 KiB + 16 B per row of the
// grid, a 16 MiB + 1 KiB part, so grids with ny*nz up to 2^20 rows).  Each
// device holds its own copy of the array until the process ends; its device
// address is looked up once per device with hipGetSymbolAddress and only
// that address is used from host code.  Every call on a device shares that
// device's array, so calls that overlap on one device race; each call
// rewrites the flags and the table it reads before reading them (the
// comparison kernels write them from the buffers).
#include <mutex>

#include "st_internal.hpp"

namespace st {

constexpr size_t kPart2 = 4 * 1024;                // row form: shell flag
constexpr size_t kPart3 = (16u << 20) + 1024;      // streaming form: flags + x-shell table
constexpr size_t kScratchBytes = kPart2 + kPart3;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mutex;
static void *g_scratch_ptr[kMaxDevices];

// Device address of `scratch` on the current device (looked up on the first
// call per device, with that device current).
static st_status scratch_address(void **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_check(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(ST_EDEVICE, "device %d beyond the scratch table", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_check(e, "hipGetSymbolAddress(scratch)");
        if (!ptr) return fail(ST_ELAUNCH, "hipGetSymbolAddress(scratch) returned null");
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return ST_OK;
}

}  // namespace st

extern "C" st_status stencil_dispatch_noalloc(const st_params *p, void *d_u1, void *d_u2, void *stream,
                                              int *result_in_u2)
{
    using namespace st;
    void *base = nullptr;
    st_status s = scratch_address(&base);
    if (s != ST_OK) return s;
    // slots 0 and 1 (residual partials, batch tables) serve other entry
    // points and get no part here
    const WsFixed fixed{base, {0, 0, 0, kPart2}, {0, 0, kPart2, kPart3}};
    workspace_use_fixed(&fixed);
    // no elapsed time: timing would create events
    s = stencil_dispatch(p, d_u1, d_u2, stream, result_in_u2, nullptr);
    workspace_use_fixed(nullptr);
    return s;
}
