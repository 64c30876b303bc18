// Internal host-side types shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gis.h"
#include "kernels.h"

namespace gis {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct HipError {
    hipError_t e;
    std::string what;
};

#define HIP_CHECK(call)                                  \
    do {                                                 \
        hipError_t e_ = (call);                          \
        if (e_ != hipSuccess) throw ::gis::HipError{e_, #call}; \
    } while (0)

// grow-only device allocation
template <typename T>
T *ensure(DevBuf &b, size_t n) {
    const size_t need = std::max<size_t>(n, 1) * sizeof(T);
    if (b.bytes < need) {
        const size_t alloc = std::max(need, b.bytes + b.bytes / 2);
        if (b.p) HIP_CHECK(hipFree(b.p));
        b.p = nullptr;
        b.bytes = 0;
        const hipError_t e = hipMalloc(&b.p, alloc);
        if (e != hipSuccess) {
            b.p = nullptr;
            throw HipError{e, "hipMalloc of " + std::to_string(alloc) + " bytes"};
        }
        b.bytes = alloc;
    }
    return static_cast<T *>(b.p);
}

inline void release(DevBuf &b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

struct DevIndexStore {
    DevBuf rs, starts, lines;  // lines: DevIndex::lines, built from rs and starts
    uint64_t n = 0;
    int bits = 0;
    uint32_t filter_cutoff = 0;
    gis_params params{};
    uint64_t distinct = 0, once = 0, high = 0, mid = 0, index_cutoff = 0;
    bool loaded = false;
    // bumped whenever the table is dropped: launch sequences captured with
    // the previous table's addresses and parameters are stale from then on
    uint64_t epoch = 0;
};

// free the resident table (before another is loaded)
inline void drop_table(DevIndexStore &ix) {
    ix.loaded = false;
    release(ix.rs);
    release(ix.starts);
    release(ix.lines);
    ix.epoch++;
}

inline int fail(gis_ctx *ctx, int code, const std::string &msg);

// device buffers and results of the pair step (pairs.cpp)
struct PairWork;
void pairwork_free(PairWork *w);
struct DmaEngines;

}  // namespace gis

struct gis_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::recursive_mutex mu;
    std::string err;
    gis::DevIndexStore ix;
    gis_batch *scratch = nullptr;  // batch reused by the host-buffer entry points
    int n_cu = 256;                // compute units of the device (persistent grids)
    gis_fused_config fcfg{4, 1, 0, 0, 8};
    gis::PairWork *pairs = nullptr;  // the pair step on NAM lists from host memory
    // lanes of gis_find_nams_stream: batches with a stream and a workspace
    // each, kept so that later calls reuse their buffers and graphs
    std::vector<gis_batch *> lanes;
    // the copy engines those lanes drive directly (stream.cpp), found on first use
    gis::DmaEngines *dma = nullptr;
    // concurrent gis_find_nams_qrs calls waiting to be run together
    std::mutex comb_mu;
    std::condition_variable comb_cv;
    std::vector<void *> comb_queue;
    bool comb_busy = false;
};

namespace gis {

inline int fail(gis_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return code;
}

inline int hip_fail(gis_ctx *ctx, const HipError &e) {
    return fail(ctx, GIS_ERR_HIP, e.what + ": " + hipGetErrorString(e.e));
}

inline DevIndex dev_index(const gis_ctx *ctx) {
    DevIndex d;
    d.rs = static_cast<const uint4 *>(ctx->ix.rs.p);
    d.starts = static_cast<const uint64_t *>(ctx->ix.starts.p);
    d.lines = static_cast<const uint4 *>(ctx->ix.lines.p);
    d.n = ctx->ix.n;
    d.bits = ctx->ix.bits;
    d.filter_cutoff = ctx->ix.filter_cutoff;
    d.k = ctx->ix.params.k;
    return d;
}

// the batch machinery (batch.cpp)
void batch_free(gis_batch *b);
// the stream lanes' copy engines (stream.cpp)
void dma_release(DmaEngines *e);
// the allocation-free lookup's device array on `device`, the current device
// (noalloc.hip); once per device
int noalloc_setup(int device);

}  // namespace gis

#define API_BEGIN(ctx)                                                 \
    if (!(ctx)) return GIS_ERR_BAD_PARAMETER;                          \
    std::lock_guard<std::recursive_mutex> lock_((ctx)->mu);            \
    try {                                                              \
        HIP_CHECK(hipSetDevice((ctx)->device));
#define API_END(ctx)                                                   \
    }                                                                  \
    catch (const ::gis::HipError &e) { return ::gis::hip_fail((ctx), e); } \
    catch (const std::bad_alloc &) { return ::gis::fail((ctx), GIS_ERR_NOMEM, "host allocation failed"); }
