// C ABI: the bucket lookup (gis_lookup_batch) for callers that cannot
// allocate device memory. Its device staging -- the hashes copied in and
// the positions the kernel writes -- lives in one file-scope device array
// instead of the context's workspace. Each GPU holds its own copy of the
// array for the life of the process; its device address is looked up once
// per device when a context is created on that device, and only that
// address is handed to copies and kernels.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "driver_internal.h"

namespace {

// hashes at [0, GIS_NOALLOC_LOOKUP_MAX), positions at [GIS_NOALLOC_LOOKUP_MAX, 2 GIS_NOALLOC_LOOKUP_MAX)
static __device__ __align__(256) uint64_t scratch[2 * GIS_NOALLOC_LOOKUP_MAX];

constexpr int MAX_DEVICES = 64;
std::mutex setup_mu;
uint64_t *scratch_ptr[MAX_DEVICES] = {};  // device address of each GPU's array

}  // namespace

namespace gis {

// Look up the array's device address on `device` (the current device) once.
int noalloc_setup(int device) {
    if (device < 0 || device >= MAX_DEVICES) return GIS_ERR_BAD_PARAMETER;
    std::lock_guard<std::mutex> lk(setup_mu);
    if (scratch_ptr[device]) return GIS_OK;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return GIS_ERR_HIP;
    scratch_ptr[device] = static_cast<uint64_t *>(p);
    return GIS_OK;
}

}  // namespace gis

using namespace gis;

extern "C" int gis_lookup_batch_noalloc(gis_ctx *ctx, const uint64_t *hashes, uint64_t n, uint64_t *positions) {
    API_BEGIN(ctx)
    if (!ctx->ix.loaded) return fail(ctx, GIS_ERR_STATE, "no index loaded");
    if (n && (!hashes || !positions)) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null argument");
    if (n > GIS_NOALLOC_LOOKUP_MAX)
        return fail(ctx, GIS_ERR_CAPACITY,
                    std::to_string(n) + " hashes exceed the allocation-free lookup's maximum of " +
                        std::to_string((uint64_t)GIS_NOALLOC_LOOKUP_MAX));
    if (n == 0) return GIS_OK;
    uint64_t *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(setup_mu);
        if (ctx->device >= 0 && ctx->device < MAX_DEVICES) base = scratch_ptr[ctx->device];
    }
    if (!base) return fail(ctx, GIS_ERR_HIP, "the allocation-free lookup's device array was not set up");
    uint64_t *dh = base, *dp = base + GIS_NOALLOC_LOOKUP_MAX;
    HIP_CHECK(hipMemcpyAsync(dh, hashes, 8 * n, hipMemcpyHostToDevice, ctx->stream));
    launch_find_flat(dev_index(ctx), dh, n, dp, ctx->stream);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(positions, dp, 8 * n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return GIS_OK;
    API_END(ctx)
}
