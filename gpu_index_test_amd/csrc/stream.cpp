// C ABI: host-to-host mapping of many reads, the reference GPU driver's
// per-batch loop (step2/main.cu:4396-4729: pack a batch, copy it to the
// device, run the stages, hand the NAMs to the host) with the batches of
// several lanes in flight. A lane is a batch with a stream and a workspace of
// its own, driven by a host thread of its own: it derives the batch's
// per-read arrays, enqueues the copies of the reads to HBM and the run (a
// replayed graph), waits for the run, and copies the NAM arrays back into the
// caller's output. While one lane's copies cross PCIe another lane's kernels
// run, so the links and the compute units stay busy together.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "batch_internal.h"

using namespace gis;

namespace {

struct StreamJob {
    gis_ctx *ctx = nullptr;
    const char *seqs = nullptr;
    const uint64_t *offs = nullptr;
    const uint32_t *lens = nullptr;
    uint64_t n = 0, nb = 0;
    uint32_t per = 0;
    gis_map_options opt{};
    gis_nam_out *out = nullptr;
    double *batch_ms = nullptr;
    std::mutex m;
    std::condition_variable cv;
    // NAMs land in read order: batch i's NAMs start after those of batches
    // 0 .. i-1, known once batch i-1 has counted its own
    uint64_t next = 0;  // the batch whose first NAM's place is `base`
    uint64_t base = 0;
    bool failed = false;
    int code = GIS_OK;
    std::string msg;
};

void fail_job(StreamJob &J, int code, const std::string &msg) {
    {
        std::lock_guard<std::mutex> lk(J.m);
        if (!J.failed) {
            J.failed = true;
            J.code = code;
            J.msg = msg;
        }
    }
    J.cv.notify_all();
}

// Lane `lane` of `lanes` maps batches lane, lane + lanes, ... in order.
void lane_main(StreamJob &J, gis_batch *b, uint32_t lane, uint32_t lanes) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    try {
        HIP_CHECK(hipSetDevice(J.ctx->device));
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        gis_nam_out *o = J.out;
        for (uint64_t i = lane; i < J.nb; i += lanes) {
            {
                std::lock_guard<std::mutex> lk(J.m);
                if (J.failed) break;
            }
            const uint64_t lo = i * J.per, hi = std::min(J.n, lo + J.per);
            const uint32_t k = (uint32_t)(hi - lo);
            uint64_t base = UINT64_MAX;
            for (uint64_t r = lo; r < hi; r++) base = std::min(base, J.offs[r]);
            // e0 is recorded just before the batch's first copy to HBM
            if (batch_stage(b, J.seqs, J.offs + lo, J.lens + lo, k, false, base, e0) != GIS_OK)
                throw HipError{hipErrorInvalidValue, "staging batch " + std::to_string(i) + " failed"};
            const RunOpts r = run_opts(b, J.opt, 0, STOP_NONE);
            batch_launch(b, r);
            batch_complete(b, r);  // regrowth reruns included
            const uint64_t tot = b->ws->tot_nams;
            uint64_t mine = 0;
            {
                std::unique_lock<std::mutex> lk(J.m);
                J.cv.wait(lk, [&] { return J.failed || J.next == i; });
                if (J.failed) break;
                mine = J.base;
                J.base += tot;
                J.next = i + 1;
            }
            J.cv.notify_all();
            hipStream_t st = b->stream;
            HIP_CHECK(hipMemcpyAsync(o->nam_offsets + lo, b->ws->nam_off.p, 8ull * k, hipMemcpyDeviceToHost, st));
            if (tot && o->nams && mine + tot <= o->nam_capacity)
                HIP_CHECK(hipMemcpyAsync(o->nams + mine, b->ws->nams_out.p, sizeof(gis_nam) * tot,
                                         hipMemcpyDeviceToHost, st));
            if (o->nonrepetitive_fraction)
                HIP_CHECK(hipMemcpyAsync(o->nonrepetitive_fraction + lo, b->ws->frac.p, 4ull * k,
                                         hipMemcpyDeviceToHost, st));
            if (o->flags) HIP_CHECK(hipMemcpyAsync(o->flags + lo, b->ws->flags.p, k, hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipEventRecord(e1, st));
            HIP_CHECK(hipEventSynchronize(e1));
            for (uint64_t x = lo; x < hi; x++) o->nam_offsets[x] += mine;
            if (J.batch_ms) {
                float ms = 0;
                HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                J.batch_ms[i] = ms;
            }
        }
    } catch (const HipError &e) {
        fail_job(J, GIS_ERR_HIP, e.what + ": " + hipGetErrorString(e.e));
        if (b->stream) (void)hipStreamSynchronize(b->stream);
    } catch (const std::bad_alloc &) {
        fail_job(J, GIS_ERR_NOMEM, "host allocation failed");
        if (b->stream) (void)hipStreamSynchronize(b->stream);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
}

}  // namespace

extern "C" int gis_find_nams_stream(gis_ctx *ctx, const char *seqs, const uint64_t *offsets, const uint32_t *lens,
                                    uint64_t n_reads, uint32_t batch_reads, uint32_t lanes,
                                    const gis_map_options *opt, gis_nam_out *out, double *batch_ms) {
    API_BEGIN(ctx)
    if (!ctx->ix.loaded) return fail(ctx, GIS_ERR_STATE, "no index loaded");
    if (!out || !out->nam_offsets) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null output");
    if (n_reads && (!seqs || !offsets || !lens)) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null read arrays");
    if (batch_reads < 1 || batch_reads > (1u << 24))
        return fail(ctx, GIS_ERR_BAD_PARAMETER, "batch_reads must be in 1..16777216");
    if (lanes < 1 || lanes > 16) return fail(ctx, GIS_ERR_BAD_PARAMETER, "lanes must be in 1..16");
    for (uint64_t r = 0; r < n_reads; r++)
        if (lens[r] > GIS_MAX_READ_LEN)
            return fail(ctx, GIS_ERR_BAD_PARAMETER,
                        "read " + std::to_string(r) + " longer than " + std::to_string(GIS_MAX_READ_LEN));
    out->n_nams = 0;
    out->nam_offsets[0] = 0;
    if (n_reads == 0) return GIS_OK;
    StreamJob J;
    J.ctx = ctx;
    J.seqs = seqs;
    J.offs = offsets;
    J.lens = lens;
    J.n = n_reads;
    J.per = batch_reads;
    J.nb = (n_reads + batch_reads - 1) / batch_reads;
    J.opt = opts_or_default(opt);
    J.out = out;
    J.batch_ms = batch_ms;
    const uint32_t L = (uint32_t)std::min<uint64_t>(lanes, J.nb);
    while (ctx->lanes.size() < L) ctx->lanes.push_back(batch_new_stream(ctx));
    std::vector<std::thread> th;
    try {
        for (uint32_t j = 0; j < L; j++) th.emplace_back(lane_main, std::ref(J), ctx->lanes[j], j, L);
    } catch (const std::system_error &) {
        fail_job(J, GIS_ERR_NOMEM, "cannot start a lane thread");
    }
    for (auto &t : th) t.join();
    if (J.failed) return fail(ctx, J.code, J.msg);
    out->nam_offsets[n_reads] = J.base;
    out->n_nams = J.base;
    if (J.base > out->nam_capacity || (J.base && !out->nams))
        return fail(ctx, GIS_ERR_CAPACITY,
                    "NAM capacity " + std::to_string(out->nam_capacity) + " < " + std::to_string(J.base));
    return GIS_OK;
    API_END(ctx)
}

extern "C" int gis_host_alloc(uint64_t bytes, void **p) {
    if (!p) return GIS_ERR_BAD_PARAMETER;
    *p = nullptr;
    if (hipHostMalloc(p, std::max<uint64_t>(bytes, 1)) != hipSuccess) {
        *p = nullptr;
        return GIS_ERR_NOMEM;
    }
    return GIS_OK;
}

extern "C" void gis_host_free(void *p) {
    if (p) (void)hipHostFree(p);
}
