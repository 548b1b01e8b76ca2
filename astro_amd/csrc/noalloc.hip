// noalloc.hip -- an exclusive scan that allocates nothing, for callers that cannot allocate
// device memory (code under graph capture, tight serving loops, harnesses that forbid it).
//
// astro_scan_noalloc runs the kernels of astro_scan (k_sort.hip: per-tile sums, then each
// tile adds up the sums before it and scans itself) with the per-tile sums in a file-scope
// device array instead of a context's scratch buffer.  No context is needed.  Each GPU holds
// its own copy of the array (kNoAllocScanTiles words, 64 KiB) until the process ends; its
// address is looked up once per device, on the first call there.  The sums are rewritten in
// full by every call, so nothing needs clearing.  Calls that overlap on one GPU share the
// array and race: serialise them (one stream, or events).
#include <hip/hip_runtime.h>

#include <mutex>

#include "launch.hpp"

using namespace astro;

namespace {
constexpr uint32_t kNoAllocScanTiles = 16384;  // scans of up to 16,384 x 2,048 values
constexpr int kMaxDevices = 64;

// The per-tile sums.  In host code this name is only a handle for hipGetSymbolAddress: the
// kernels receive the device address that call returns, never the name.
static __device__ __align__(256) uint32_t scratch[kNoAllocScanTiles];

std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[kMaxDevices] = {};

// The scratch array's address on `device` (the current device), looked up on first use.
int scratch_on(int device, uint32_t** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
            return ASTRO_ERR_HIP;
        g_scratch_ptr[device] = static_cast<uint32_t*>(p);
    }
    *out = g_scratch_ptr[device];
    return ASTRO_OK;
}
}  // namespace

extern "C" int astro_scan_noalloc(const uint32_t* d_in, uint32_t* d_out, uint32_t len,
                                  void* stream) {
    if ((!d_in || !d_out) && len) return ASTRO_ERR_ARG;
    if (len == 0) return ASTRO_OK;
    if (scan_tiles(len) > kNoAllocScanTiles) return ASTRO_ERR_CAPACITY;
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return ASTRO_ERR_HIP;
    if (device < 0 || device >= kMaxDevices) return ASTRO_ERR_ARG;
    uint32_t* sums = nullptr;
    if (int rc = scratch_on(device, &sums)) return rc;
    launch_scan(d_in, d_out, len, sums, static_cast<hipStream_t>(stream));
    return hipGetLastError() == hipSuccess ? ASTRO_OK : ASTRO_ERR_HIP;
}
