// C ABI: host-to-host mapping of many reads, the reference GPU driver's
// per-batch loop (step2/main.cu:4396-4729: pack a batch, copy it to the
// device, run the stages, hand the NAMs to the host) with the batches of
// several lanes in flight. A lane is a batch with a stream and a workspace of
// its own, driven by a host thread of its own: it derives the batch's
// per-read arrays, copies the reads to HBM, runs the batch (a replayed graph),
// and copies the NAM arrays back into the caller's output. While one lane's
// copies cross PCIe another lane's kernels run, so the links and the compute
// units stay busy together.
//
// The copies run on two SDMA engines of the GPU, one per direction, driven
// through ROCr. HIP puts both directions of its own copies on one engine, so
// a batch's reads going in wait for another batch's NAMs coming out; on two
// engines the PCIe link carries both directions at once. Host memory that no
// engine can reach (pageable memory) is copied through HIP instead.
#include <hip/hip_runtime.h>
#include <hsa/hsa.h>
#include <hsa/hsa_ext_amd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "batch_internal.h"

namespace gis {

struct DmaEngines {
    std::atomic<bool> ok{false};  // engines found and not refusing copies
    bool hsa_up = false;          // hsa_init succeeded (balanced by hsa_shut_down)
    hsa_agent_t gpu{}, cpu{};
    uint32_t h2d = 0, d2h = 0;  // hsa_amd_sdma_engine_id_t bits
    std::atomic<uint64_t> engine_copies{0}, hip_copies{0}, refused{0};
};

void dma_release(DmaEngines *e) {
    if (!e) return;
    if (e->hsa_up) (void)hsa_shut_down();
    delete e;
}

}  // namespace gis

using namespace gis;

namespace {

struct AgentSearch {
    uint32_t domain = 0, bdf = 0;
    hsa_agent_t gpu{}, cpu{};
    bool have_gpu = false, have_cpu = false;
};

hsa_status_t find_agents(hsa_agent_t a, void *arg) {
    auto *s = static_cast<AgentSearch *>(arg);
    hsa_device_type_t t;
    if (hsa_agent_get_info(a, HSA_AGENT_INFO_DEVICE, &t) != HSA_STATUS_SUCCESS) return HSA_STATUS_SUCCESS;
    if (t == HSA_DEVICE_TYPE_CPU && !s->have_cpu) {
        s->cpu = a;
        s->have_cpu = true;
    }
    if (t == HSA_DEVICE_TYPE_GPU && !s->have_gpu) {
        uint32_t bdf = 0, dom = 0;
        if (hsa_agent_get_info(a, (hsa_agent_info_t)HSA_AMD_AGENT_INFO_BDFID, &bdf) == HSA_STATUS_SUCCESS &&
            hsa_agent_get_info(a, (hsa_agent_info_t)HSA_AMD_AGENT_INFO_DOMAIN, &dom) == HSA_STATUS_SUCCESS &&
            bdf == s->bdf && dom == s->domain) {
            s->gpu = a;
            s->have_gpu = true;
        }
    }
    return HSA_STATUS_SUCCESS;
}

uint32_t lowest_bit(uint32_t m) { return m & (~m + 1); }
uint32_t highest_bit(uint32_t m) { return m ? 1u << (31 - __builtin_clz(m)) : 0u; }

// The ROCr agent of HIP device `device` (matched by PCI location) and two of
// its SDMA engines: for host-to-device copies the lowest engine ROCr prefers
// for that direction, for device-to-host the highest preferred one that is
// not that engine. Without them every copy goes through HIP.
DmaEngines *dma_setup(int device) {
    auto *E = new DmaEngines();
    char pci[32] = {};
    unsigned dom = 0, bus = 0, dev = 0, fn = 0;
    if (hipDeviceGetPCIBusId(pci, sizeof(pci), device) != hipSuccess ||
        std::sscanf(pci, "%x:%x:%x.%x", &dom, &bus, &dev, &fn) != 4)
        return E;
    if (hsa_init() != HSA_STATUS_SUCCESS) return E;
    E->hsa_up = true;
    AgentSearch s;
    s.domain = dom;
    s.bdf = (bus << 8) | (dev << 3) | fn;
    if (hsa_iterate_agents(find_agents, &s) != HSA_STATUS_SUCCESS || !s.have_gpu || !s.have_cpu) return E;
    uint32_t st_h = 0, st_d = 0, pr_h = 0, pr_d = 0;
    if (hsa_amd_memory_copy_engine_status(s.gpu, s.cpu, &st_h) != HSA_STATUS_SUCCESS ||
        hsa_amd_memory_copy_engine_status(s.cpu, s.gpu, &st_d) != HSA_STATUS_SUCCESS)
        return E;
    if (hsa_amd_memory_get_preferred_copy_engine(s.gpu, s.cpu, &pr_h) != HSA_STATUS_SUCCESS) pr_h = 0;
    if (hsa_amd_memory_get_preferred_copy_engine(s.cpu, s.gpu, &pr_d) != HSA_STATUS_SUCCESS) pr_d = 0;
    uint32_t h = lowest_bit(pr_h & st_h);
    if (!h) h = lowest_bit(st_h);
    const uint32_t dm = st_d & ~h;
    uint32_t d = highest_bit(pr_d & dm);
    if (!d) d = lowest_bit(dm);
    if (!h || !d) return E;
    E->gpu = s.gpu;
    E->cpu = s.cpu;
    E->h2d = h;
    E->d2h = d;
    E->ok = true;
    return E;
}

// [p, p + bytes) lies in one allocation that the GPU reaches at the same
// address (pinned memory from hipHostMalloc), so an engine can copy it.
bool engine_reachable(const void *p, size_t bytes) {
    hsa_amd_pointer_info_t info{};
    info.size = sizeof(info);
    if (hsa_amd_pointer_info(p, &info, nullptr, nullptr, nullptr) != HSA_STATUS_SUCCESS) return false;
    if (info.type != HSA_EXT_POINTER_TYPE_HSA || !info.hostBaseAddress ||
        info.agentBaseAddress != info.hostBaseAddress)
        return false;
    const char *base = static_cast<const char *>(info.hostBaseAddress), *q = static_cast<const char *>(p);
    return q >= base && bytes <= info.sizeInBytes && (size_t)(q - base) <= info.sizeInBytes - bytes;
}

// Runs `items`, all in one direction. Those whose host side an engine can
// reach go to that direction's engine and are waited for on `sig`; the
// others are enqueued through HIP on `st` (host-to-device copies then
// precede the run on the stream; device-to-host copies are waited for here).
void run_copies(DmaEngines *E, bool to_device, const std::vector<CopyItem> &items, hsa_signal_t sig,
                bool have_sig, hipStream_t st) {
    const hipMemcpyKind kind = to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
    std::vector<const CopyItem *> eng, hip;
    for (const CopyItem &c : items) {
        if (!c.bytes) continue;
        const bool reach = have_sig && E->ok.load() && engine_reachable(to_device ? c.src : c.dst, c.bytes);
        (reach ? eng : hip).push_back(&c);
    }
    size_t issued = 0;
    if (!eng.empty()) {
        hsa_signal_store_screlease(sig, (hsa_signal_value_t)eng.size());
        for (const CopyItem *c : eng) {
            const hsa_status_t rc =
                to_device ? hsa_amd_memory_async_copy_on_engine(c->dst, E->gpu, c->src, E->cpu, c->bytes, 0, nullptr,
                                                                sig, (hsa_amd_sdma_engine_id_t)E->h2d, true)
                          : hsa_amd_memory_async_copy_on_engine(c->dst, E->cpu, c->src, E->gpu, c->bytes, 0, nullptr,
                                                                sig, (hsa_amd_sdma_engine_id_t)E->d2h, true);
            if (rc != HSA_STATUS_SUCCESS) break;
            issued++;
        }
        if (issued < eng.size()) {
            // an engine that refuses a copy is not used again by this context
            E->ok = false;
            E->refused += eng.size() - issued;
            hsa_signal_subtract_screlease(sig, (hsa_signal_value_t)(eng.size() - issued));
            hip.insert(hip.end(), eng.begin() + issued, eng.end());
        }
        E->engine_copies += issued;
    }
    for (const CopyItem *c : hip) HIP_CHECK(hipMemcpyAsync(c->dst, c->src, c->bytes, kind, st));
    E->hip_copies += hip.size();
    if (issued) hsa_signal_wait_scacquire(sig, HSA_SIGNAL_CONDITION_EQ, 0, UINT64_MAX, HSA_WAIT_STATE_BLOCKED);
    if (!to_device && !hip.empty()) HIP_CHECK(hipStreamSynchronize(st));
}

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
    DmaEngines *E = J.ctx->dma;
    hsa_signal_t sig{};
    const bool have_sig = E->ok.load() && hsa_signal_create(0, 0, nullptr, &sig) == HSA_STATUS_SUCCESS;
    try {
        HIP_CHECK(hipSetDevice(J.ctx->device));
        gis_nam_out *o = J.out;
        std::vector<CopyItem> in, back;
        for (uint64_t i = lane; i < J.nb; i += lanes) {
            {
                std::lock_guard<std::mutex> lk(J.m);
                if (J.failed) break;
            }
            const auto t0 = std::chrono::steady_clock::now();
            const uint64_t lo = i * J.per, hi = std::min(J.n, lo + J.per);
            const uint32_t k = (uint32_t)(hi - lo);
            uint64_t base = UINT64_MAX;
            for (uint64_t r = lo; r < hi; r++) base = std::min(base, J.offs[r]);
            in.clear();
            if (batch_stage(b, J.seqs, J.offs + lo, J.lens + lo, k, false, base, nullptr, &in,
                            J.opt.mode == GIS_MODE_FUSED) != GIS_OK)
                throw HipError{hipErrorInvalidValue, "staging batch " + std::to_string(i) + " failed"};
            hipStream_t st = b->stream;
            run_copies(E, true, in, sig, have_sig, st);
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
            back.clear();
            back.push_back({o->nam_offsets + lo, b->ws->nam_off.p, 8ull * k});
            if (tot && o->nams && mine + tot <= o->nam_capacity)
                back.push_back({o->nams + mine, b->ws->nams_out.p, sizeof(gis_nam) * tot});
            if (o->nonrepetitive_fraction) back.push_back({o->nonrepetitive_fraction + lo, b->ws->frac.p, 4ull * k});
            if (o->flags) back.push_back({o->flags + lo, b->ws->flags.p, (size_t)k});
            run_copies(E, false, back, sig, have_sig, st);
            for (uint64_t x = lo; x < hi; x++) o->nam_offsets[x] += mine;
            if (J.batch_ms)
                J.batch_ms[i] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
    } catch (const HipError &e) {
        fail_job(J, GIS_ERR_HIP, e.what + ": " + hipGetErrorString(e.e));
        if (b->stream) (void)hipStreamSynchronize(b->stream);
    } catch (const std::bad_alloc &) {
        fail_job(J, GIS_ERR_NOMEM, "host allocation failed");
        if (b->stream) (void)hipStreamSynchronize(b->stream);
    }
    if (have_sig) (void)hsa_signal_destroy(sig);
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
    if (!ctx->dma) ctx->dma = dma_setup(ctx->device);
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

extern "C" int gis_stream_copy_stats(gis_ctx *ctx, gis_copy_stats *s) {
    API_BEGIN(ctx)
    if (!s) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null output");
    *s = gis_copy_stats{};
    if (DmaEngines *E = ctx->dma) {
        s->h2d_engine = E->ok.load() ? E->h2d : 0;
        s->d2h_engine = E->ok.load() ? E->d2h : 0;
        s->engine_copies = E->engine_copies.load();
        s->hip_copies = E->hip_copies.load();
        s->refused = E->refused.load();
    }
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
