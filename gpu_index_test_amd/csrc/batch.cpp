// C ABI: device batches and the whole-path / stage entry points.
//
// A batch holds reads resident in HBM and every buffer of the path. A run is
// a fixed sequence of launches on the context's stream with no host round
// trip: variable-size intermediates (hits, rescue hits) are placed by device
// scans of per-read counts and checked against their capacity on the device,
// and the single synchronisation at the end reads the totals and the
// overflow bits. When an array was too small it grows to the total the scan
// reported and the batch runs again; results never depend on it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <mutex>
#include <vector>

#include "batch_internal.h"

using namespace gis;

namespace gis {

gis_batch *batch_new(gis_ctx *ctx, Workspace *share) {
    gis_batch *b = new gis_batch();
    b->ctx = ctx;
    b->stream = ctx->stream;
    b->ws = share ? share : new Workspace();
    b->ws->refs++;
    return b;
}

gis_batch *batch_new_stream(gis_ctx *ctx) {
    gis_batch *b = batch_new(ctx, nullptr);
    // default priority, like the context's stream; the workspace owns it, so
    // batches sharing the workspace keep it while any of them lives
    if (hipStreamCreateWithFlags(&b->ws->stream, hipStreamNonBlocking) != hipSuccess) {
        b->ws->stream = nullptr;
        batch_free(b);
        throw HipError{hipErrorOutOfMemory, "hipStreamCreateWithFlags"};
    }
    b->stream = b->ws->stream;
    return b;
}

// the device arrays of a workspace, and of a batch's own inputs
static std::vector<DevBuf *> ws_bufs(Workspace *w) {
    return {&w->long_scratch, &w->rs_cnt, &w->rs_nfwd, &w->rs_hash, &w->rs_span,
            &w->sy_cnt, &w->sy_hash, &w->sy_pos, &w->pos, &w->hcnt, &w->rcnt, &w->lk_bytes, &w->lk_stream,
            &w->hit_total, &w->rhit_total, &w->frac, &w->flags, &w->nam_cnt, &w->hit_off,
            &w->rhit_off, &w->hits, &w->rhits, &w->keys, &w->rkeys, &w->nam_ws, &w->rnam_ws,
            &w->nam_fin, &w->rnam_fin, &w->nam_slots, &w->nam_off, &w->nams_out, &w->scan_tmp,
            &w->ctl, &w->spill_ids, &w->gwork, &w->split_list, &w->sel_ids};
}

static std::vector<DevBuf *> input_bufs(gis_batch *b) {
    return {&b->seqs, &b->read_off, &b->read_len, &b->rs_off, &b->sy_off, &b->long_off, &b->queue};
}

static void workspace_release(Workspace *w) {
    for (DevBuf *x : ws_bufs(w)) release(*x);
    if (w->h_ctl) (void)hipHostFree(w->h_ctl);
    pairwork_free(w->pairs);
    if (w->stream) (void)hipStreamDestroy(w->stream);
    delete w;
}

uint64_t batch_device_bytes(gis_batch *b) {
    uint64_t t = 0;
    for (DevBuf *x : input_bufs(b)) t += x->bytes;
    for (DevBuf *x : ws_bufs(b->ws)) t += x->bytes;
    return t;
}

void batch_trim(gis_batch *b) {
    HIP_CHECK(hipStreamSynchronize(b->stream));
    for (auto &g : b->graphs) {
        HIP_CHECK(hipGraphExecDestroy(g.exec));
        HIP_CHECK(hipGraphDestroy(g.graph));
    }
    b->graphs.clear();
    for (DevBuf *x : input_bufs(b)) release(*x);
    for (DevBuf *x : ws_bufs(b->ws)) release(*x);
    b->ws->hit_cap = b->ws->rhit_cap = b->ws->out_cap = 0;
    b->ws->merge_scratch = false;
    b->ws->owner = nullptr;
    b->ran = false;
    b->n_reads = 0;
}

void batch_free(gis_batch *b) {
    if (!b) return;
    for (DevBuf *x : input_bufs(b)) release(*x);
    for (auto &e : b->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &g : b->graphs) {
        (void)hipGraphExecDestroy(g.exec);
        (void)hipGraphDestroy(g.graph);
    }
    if (b->ws) {
        if (b->ws->owner == b) b->ws->owner = nullptr;
        if (b->ws->in_flight == b) b->ws->in_flight = nullptr;
        if (--b->ws->refs == 0) workspace_release(b->ws);
    }
    delete b;
}

}  // namespace gis

namespace gis {

// reads up to this length form the fused kernel's short bin
constexpr int SHORT_BIN = 160;

// long-read scratch bytes of the seeding stage for one read of length L
// (layout of long_bufs in kernels.hip)
uint64_t long_bytes(uint64_t L) {
    auto a8 = [](uint64_t x) { return (x + 7) & ~7ULL; };
    return 16 * L + 8 * (L / 64 + 2) + a8(4 * L) + a8(L) + a8(L) + 64;
}

DevBatch make_devbatch(gis_batch *b) {
    DevBatch d{};
    d.seqs = static_cast<const char *>(b->seqs.p);
    d.read_off = static_cast<const uint64_t *>(b->read_off.p);
    d.read_len = static_cast<const uint32_t *>(b->read_len.p);
    d.n_reads = b->n_reads;
    d.rs_off = static_cast<const uint64_t *>(b->rs_off.p);
    d.rs_cnt = static_cast<uint32_t *>(b->ws->rs_cnt.p);
    d.rs_nfwd = static_cast<uint32_t *>(b->ws->rs_nfwd.p);
    d.rs_hash = static_cast<uint64_t *>(b->ws->rs_hash.p);
    d.rs_span = static_cast<uint32_t *>(b->ws->rs_span.p);
    d.sy_off = static_cast<const uint64_t *>(b->sy_off.p);
    d.sy_cnt = static_cast<uint32_t *>(b->ws->sy_cnt.p);
    d.sy_hash = static_cast<uint64_t *>(b->ws->sy_hash.p);
    d.sy_pos = static_cast<uint64_t *>(b->ws->sy_pos.p);
    d.pos = static_cast<uint64_t *>(b->ws->pos.p);
    d.hcnt = static_cast<uint32_t *>(b->ws->hcnt.p);
    d.rcnt = static_cast<uint32_t *>(b->ws->rcnt.p);
    d.lk_bytes = static_cast<uint32_t *>(b->ws->lk_bytes.p);
    d.lk_stream = static_cast<uint32_t *>(b->ws->lk_stream.p);
    d.hit_total = static_cast<uint32_t *>(b->ws->hit_total.p);
    d.rhit_total = static_cast<uint32_t *>(b->ws->rhit_total.p);
    d.frac = static_cast<float *>(b->ws->frac.p);
    d.flags = static_cast<uint8_t *>(b->ws->flags.p);
    d.nam_cnt = static_cast<uint32_t *>(b->ws->nam_cnt.p);
    d.hit_off = static_cast<uint64_t *>(b->ws->hit_off.p);
    d.rhit_off = static_cast<uint64_t *>(b->ws->rhit_off.p);
    d.hits = static_cast<DevHit *>(b->ws->hits.p);
    d.rhits = static_cast<DevHit *>(b->ws->rhits.p);
    d.hit_cap = b->ws->hit_cap;
    d.rhit_cap = b->ws->rhit_cap;
    d.keys = static_cast<uint64_t *>(b->ws->keys.p);
    d.rkeys = static_cast<uint64_t *>(b->ws->rkeys.p);
    d.nam_ws = static_cast<WideNam *>(b->ws->nam_ws.p);
    d.rnam_ws = static_cast<WideNam *>(b->ws->rnam_ws.p);
    d.merge_scratch = b->ws->merge_scratch ? 1u : 0u;
    d.gwork = static_cast<uint4 *>(b->ws->gwork.p);
    d.split_list = static_cast<uint4 *>(b->ws->split_list.p);
    d.sel_ids = static_cast<uint32_t *>(b->ws->sel_ids.p);
    d.nam_fin = static_cast<gis_nam *>(b->ws->nam_fin.p);
    d.rnam_fin = static_cast<gis_nam *>(b->ws->rnam_fin.p);
    d.nam_slots = static_cast<gis_nam *>(b->ws->nam_slots.p);
    d.nam_off = static_cast<uint64_t *>(b->ws->nam_off.p);
    d.nams_out = static_cast<gis_nam *>(b->ws->nams_out.p);
    d.out_cap = b->ws->out_cap;
    d.long_off = static_cast<const uint64_t *>(b->long_off.p);
    d.long_scratch = static_cast<uint8_t *>(b->ws->long_scratch.p);
    d.ctl = static_cast<DevCtl *>(b->ws->ctl.p);
    d.spill_ids = static_cast<uint32_t *>(b->ws->spill_ids.p);
    return d;
}

// the batch's own per-read input arrays
void ensure_input_arrays(gis_batch *b, uint32_t n) {
    ensure<uint64_t>(b->read_off, n);
    ensure<uint32_t>(b->read_len, n);
    ensure<uint64_t>(b->rs_off, n + 1);
    ensure<uint64_t>(b->sy_off, n + 1);
    ensure<uint64_t>(b->long_off, n + 1);
    ensure<uint32_t>(b->queue, n);
}

// per-read arrays of the workspace
void ensure_read_arrays(gis_batch *b, uint32_t n) {
    ensure<uint32_t>(b->ws->rs_cnt, n);
    ensure<uint32_t>(b->ws->rs_nfwd, n);
    ensure<uint32_t>(b->ws->sy_cnt, n);
    ensure<uint32_t>(b->ws->lk_bytes, n);
    ensure<uint32_t>(b->ws->lk_stream, n);
    ensure<uint32_t>(b->ws->hit_total, n);
    ensure<uint32_t>(b->ws->rhit_total, n);
    ensure<uint32_t>(b->ws->nam_cnt, n);
    ensure<float>(b->ws->frac, n);
    ensure<uint8_t>(b->ws->flags, n);
    ensure<uint64_t>(b->ws->hit_off, n + 1);
    ensure<uint64_t>(b->ws->rhit_off, n + 1);
    ensure<uint64_t>(b->ws->nam_off, n + 1);
    ensure<uint64_t>(b->ws->scan_tmp, scan_tmp_words(n) + 8);
    ensure<DevCtl>(b->ws->ctl, 1);
    ensure<uint32_t>(b->ws->spill_ids, n);
    ensure<uint4>(b->ws->split_list, n);
    ensure<uint32_t>(b->ws->sel_ids, 3ull * n + 3);
    if (!b->ws->h_ctl) {
        HIP_CHECK(hipHostMalloc((void **)&b->ws->h_ctl, sizeof(DevCtl)));
        HIP_CHECK(hipHostGetDevicePointer((void **)&b->ws->h_ctl_dev, b->ws->h_ctl, 0));
    }
}

void ensure_slot_arrays(gis_batch *b, uint64_t slots) {
    ensure<uint64_t>(b->ws->rs_hash, slots);
    ensure<uint32_t>(b->ws->rs_span, slots);
    ensure<uint64_t>(b->ws->pos, slots);
    ensure<uint32_t>(b->ws->hcnt, slots);
    ensure<uint32_t>(b->ws->rcnt, slots);
}

// hit arrays, merge scratch and the NAM output (capacities in hits / NAMs)
void ensure_pools(gis_batch *b, uint64_t hit_cap, uint64_t rhit_cap, uint64_t out_cap) {
    b->ws->hit_cap = std::max(b->ws->hit_cap, hit_cap);
    b->ws->rhit_cap = std::max(b->ws->rhit_cap, rhit_cap);
    b->ws->out_cap = std::max(b->ws->out_cap, out_cap);
    ensure<DevHit>(b->ws->hits, b->ws->hit_cap);
    ensure<gis_nam>(b->ws->nam_fin, b->ws->hit_cap);
    ensure<DevHit>(b->ws->rhits, b->ws->rhit_cap);
    ensure<gis_nam>(b->ws->rnam_fin, b->ws->rhit_cap);
    if (b->ws->merge_scratch) {
        // merge scratch of reads beyond the merge's LDS budget, allocated
        // once a run has met such a read
        ensure<uint64_t>(b->ws->keys, 2 * b->ws->hit_cap);
        ensure<WideNam>(b->ws->nam_ws, b->ws->hit_cap);
        ensure<uint64_t>(b->ws->rkeys, 2 * b->ws->rhit_cap);
        ensure<WideNam>(b->ws->rnam_ws, b->ws->rhit_cap);
        ensure<uint4>(b->ws->gwork, std::max(b->ws->hit_cap, b->ws->rhit_cap) + 64);
    }
    ensure<gis_nam>(b->ws->nams_out, b->ws->out_cap);
}

void mark(gis_batch *b, const char *name) {
    if (b->n_ev >= 16) return;
    if (!b->ev[b->n_ev]) HIP_CHECK(hipEventCreate(&b->ev[b->n_ev]));
    HIP_CHECK(hipEventRecord(b->ev[b->n_ev], b->stream));
    b->ev_names.push_back(name);
    b->n_ev++;
}

// Stage reads into the batch: randstrobe-slot capacities and offsets come
// from the lengths, computed here once.
int batch_stage(gis_batch *b, const char *seqs, const uint64_t *offsets, const uint32_t *lens,
                uint32_t n, bool want_sy, uint64_t base, hipEvent_t before, std::vector<CopyItem> *defer,
                bool want_queue) {
    gis_ctx *ctx = b->ctx;
    if (!ctx->ix.loaded && !b->seed_params) return fail(ctx, GIS_ERR_STATE, "no index loaded");
    if (n > 0 && (!seqs || !offsets || !lens)) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null read arrays");
    const gis_params &p = b->seed_params ? *b->seed_params : ctx->ix.params;
    uint64_t span = 0;
    b->h_rs_off.resize(n + 1);
    b->h_sy_off.resize(n + 1);
    b->h_long_off.resize(n + 1);
    b->h_read_off.resize(n);
    b->h_rs_off[0] = b->h_sy_off[0] = b->h_long_off[0] = 0;
    b->any_long = false;
    uint64_t rs_at = 0, sy_at = 0, long_at = 0;
    uint32_t n_long = 0;
    for (uint32_t r = 0; r < n; r++) {
        const uint64_t L = lens[r];
        if (L > GIS_MAX_READ_LEN)
            return fail(ctx, GIS_ERR_BAD_PARAMETER,
                        "read " + std::to_string(r) + " longer than " + std::to_string(GIS_MAX_READ_LEN));
        if (offsets[r] < base) return fail(ctx, GIS_ERR_BAD_PARAMETER, "read offset below the staged base");
        const uint64_t o = offsets[r] - base;
        b->h_read_off[r] = o;
        span = std::max(span, o + L);
        const int64_t nk = (int64_t)L - p.k + 1;
        const int64_t nf = nk - (int64_t)p.w_min;
        const uint64_t cap_sy = nk > 0 ? (uint64_t)nk : 0;
        const uint64_t cap_rs = (nf > 0 && L >= p.w_max) ? 2 * (uint64_t)nf : 0;
        rs_at += cap_rs;
        sy_at += cap_sy;
        const bool lg = L > (uint64_t)LMAX_LDS;
        b->any_long |= lg;
        long_at += lg ? long_bytes(L) : 0;
        n_long += L > (uint64_t)SHORT_BIN ? 1u : 0u;
        b->h_rs_off[r + 1] = rs_at;
        b->h_sy_off[r + 1] = sy_at;
        b->h_long_off[r + 1] = long_at;
    }
    b->n_reads = n;
    b->want_sy = want_sy;
    b->n_long = n_long;
    b->queue_ok = want_queue;
    // the fused kernel's work queue: read ids by length, longest first
    // (counting sort over lengths capped at 1023; stable within a length)
    if (want_queue) {
        uint32_t cnt[1025] = {};
        for (uint32_t r = 0; r < n; r++) cnt[1023 - std::min<uint32_t>(lens[r], 1023u) + 1]++;
        for (int i = 1; i < 1025; i++) cnt[i] += cnt[i - 1];
        b->h_queue.resize(n);
        for (uint32_t r = 0; r < n; r++) b->h_queue[cnt[1023 - std::min<uint32_t>(lens[r], 1023u)]++] = r;
    }
    ensure_input_arrays(b, n);
    b->tot_rs_cap = b->h_rs_off[n];
    b->span = span;
    ensure<char>(b->seqs, span);
    std::vector<CopyItem> items;
    std::vector<CopyItem> &in = defer ? *defer : items;
    if (span) in.push_back({b->seqs.p, seqs + base, span});
    if (n) {
        in.push_back({b->read_off.p, b->h_read_off.data(), 8ull * n});
        in.push_back({b->read_len.p, lens, 4ull * n});
    }
    in.push_back({b->rs_off.p, b->h_rs_off.data(), 8ull * (n + 1)});
    in.push_back({b->sy_off.p, b->h_sy_off.data(), 8ull * (n + 1)});
    in.push_back({b->long_off.p, b->h_long_off.data(), 8ull * (n + 1)});
    if (n && want_queue) in.push_back({b->queue.p, b->h_queue.data(), 4ull * n});
    b->ran = false;
    if (defer) return GIS_OK;
    hipStream_t st = b->stream;
    if (before) HIP_CHECK(hipEventRecord(before, st));
    for (const CopyItem &c : items) HIP_CHECK(hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyHostToDevice, st));
    return GIS_OK;
}

int batch_upload(gis_batch *b, const char *seqs, const uint64_t *offsets, const uint32_t *lens,
                 uint32_t n, bool want_sy) {
    const int rc = batch_stage(b, seqs, offsets, lens, n, want_sy, 0, nullptr);
    if (rc) return rc;
    HIP_CHECK(hipStreamSynchronize(b->stream));
    return GIS_OK;
}

uint32_t rescue_cutoff_of(const gis_ctx *ctx, const gis_map_options &o) {
    // step2/main.cu:4204: R < 100 ? R * filter_cutoff : 100
    return o.rescue_level < 100 ? (uint32_t)(o.rescue_level * (int)ctx->ix.filter_cutoff) : 100u;
}

// The launch sequence of one run, enqueued on the context's stream.
void enqueue_run(gis_batch *b, int mode, int rescue_mode, uint32_t rcut, uint32_t hard, int stop) {
    gis_ctx *ctx = b->ctx;
    hipStream_t st = b->stream;
    const DevIndex ix = dev_index(ctx);
    const SeedParams sp = make_seed_params(b->seed_params ? *b->seed_params : ctx->ix.params);
    const DevBatch d = make_devbatch(b);
    const uint32_t n = b->n_reads;
    uint64_t *tmp = static_cast<uint64_t *>(b->ws->scan_tmp.p);
    b->n_ev = 0;
    b->ev_names.clear();
    mark(b, "start");
    launch_reset_ctl(d.ctl, st);
    ReadSet set{nullptr, nullptr, n};
    const bool fused = mode == GIS_MODE_FUSED && stop == STOP_NONE;
    if (fused) {
        const gis_fused_config &fc = ctx->fcfg;
        const unsigned grid = fc.blocks_per_cu > 0 ? (unsigned)(ctx->n_cu * fc.blocks_per_cu) : 0u;
        const uint32_t chunk = fc.queue ? (uint32_t)fc.chunk : 0u;
        const uint32_t *q = static_cast<const uint32_t *>(b->queue.p);
        bool ok;
        if (fc.length_split) {
            ok = launch_fused(d, ix, sp, rcut, hard, q, b->n_long, 320, fc.waves_per_block, grid, chunk,
                              &d.ctl->q_head[0], st) &&
                 launch_fused(d, ix, sp, rcut, hard, q + b->n_long, n - b->n_long, SHORT_BIN,
                              fc.waves_per_block, grid, chunk, &d.ctl->q_head[1], st);
        } else {
            ok = launch_fused(d, ix, sp, rcut, hard, q, n, 320, fc.waves_per_block, grid, chunk,
                              &d.ctl->q_head[0], st);
        }
        if (!ok) throw HipError{hipErrorInvalidValue, "unsupported fused configuration"};
        mark(b, "fused");
        set = ReadSet{d.spill_ids, &d.ctl->n_spill, n};
    }
    launch_seed(d, set, sp, b->want_sy ? 1 : 0, b->any_long ? 1 : 0, st);
    mark(b, "seed");
    if (stop != STOP_SEED) {
        launch_lookup(d, set, ix, rescue_mode, st);
        mark(b, "lookup");
        launch_hits(d, set, ix, st);
        mark(b, "hits");
        launch_rescue_count(d, set, ix, rcut, hard, st);
        launch_scan(d.rhit_total, d.rhit_off, n, tmp, st);
        launch_rescue_hits(d, set, ix, st);
        mark(b, "rescue");
    }
    if (stop == STOP_NONE) {
        launch_merge(d, set, 0, st);
        mark(b, "merge");
        launch_merge(d, set, 1, st);
        mark(b, "rescue_merge");
        launch_scan(d.nam_cnt, d.nam_off, n, tmp, st);
        launch_compact(d, fused ? 1 : 0, st);
        mark(b, "output");
    }
    launch_ctl_to_host(d.ctl, b->ws->h_ctl_dev, st);
}

// size the workspace for this batch's reads (grow-only; cheap when it
// already fits)
void prepare_workspace(gis_batch *b) {
    const uint32_t n = b->n_reads;
    ensure_read_arrays(b, n);
    ensure_slot_arrays(b, b->tot_rs_cap);
    if (b->want_sy) {
        ensure<uint64_t>(b->ws->sy_hash, b->h_sy_off[n]);
        ensure<uint64_t>(b->ws->sy_pos, b->h_sy_off[n]);
    }
    if (b->any_long) ensure<uint8_t>(b->ws->long_scratch, b->h_long_off[n]);
    // initial pools: about the hits of typical reads (~0.16 per base), grown
    // on demand to what a run's scans report.
    // GIS_TEST_SMALL_POOLS starts them tiny so tests exercise the overflow path.
    // the hit array holds a HIT_REGION-hit region per read, then the pool of
    // the reads with more hits
    const uint64_t regions = (uint64_t)HIT_REGION * n;
    if (const char *e = std::getenv("GIS_TEST_SMALL_POOLS"); e && *e == '1')
        ensure_pools(b, regions + 64, 64, 16);
    else
        ensure_pools(b, regions + 4ull * n + b->span / 64 + 4096, 4ull * n + 4096, 12ull * n + 1024);
}

RunOpts run_opts(const gis_batch *b, const gis_map_options &o, int rescue_mode, int stop) {
    RunOpts r;
    r.mode = o.mode;
    r.rescue_mode = rescue_mode;
    r.rcut = rescue_cutoff_of(b->ctx, o);
    r.hard = o.hard_cutoff;
    r.stop = stop;
    r.graph = stop == STOP_NONE && (o.flags & GIS_RUN_NO_GRAPH) == 0;
    return r;
}

// Enqueue one run on the batch's stream. With r.graph the launch sequence is
// captured once per configuration (buffers, mode, cutoffs, fused launch
// configuration) and replayed; a batch keeps the graphs of its last 8
// configurations.
void batch_launch(gis_batch *b, const RunOpts &r) {
    gis_ctx *ctx = b->ctx;
    hipStream_t st = b->stream;
    prepare_workspace(b);
    b->ws->owner = b;
    b->ran = false;
    if (r.mode == GIS_MODE_FUSED && r.stop == STOP_NONE) {
        if (!b->queue_ok) throw HipError{hipErrorInvalidValue, "fused run of reads staged without the work queue"};
        ensure<gis_nam>(b->ws->nam_slots, (size_t)b->n_reads * F_NAMCAP);
    }
    // graphs captured against a table that has since been dropped hold its
    // freed addresses and old parameters: never replay them
    bool stale = false;
    for (const auto &g : b->graphs) stale |= g.epoch != ctx->ix.epoch;
    if (stale) {
        HIP_CHECK(hipStreamSynchronize(st));
        for (auto it = b->graphs.begin(); it != b->graphs.end();) {
            if (it->epoch == ctx->ix.epoch) {
                ++it;
                continue;
            }
            HIP_CHECK(hipGraphExecDestroy(it->exec));
            HIP_CHECK(hipGraphDestroy(it->graph));
            it = b->graphs.erase(it);
        }
    }
    if (r.graph) {
        const DevBatch key = make_devbatch(b);
        const gis_fused_config &fc = ctx->fcfg;
        const int cfg[11] = {r.mode, r.rescue_mode, (int)r.rcut, (int)r.hard, r.stop,
                             (b->want_sy ? 1 : 0) | (b->any_long ? 2 : 0), fc.waves_per_block,
                             fc.length_split, fc.chunk, fc.blocks_per_cu, (int)b->n_long};
        gis_batch::GraphEntry *hit = nullptr;
        for (auto &g : b->graphs)
            if (std::memcmp(&key, &g.key, sizeof key) == 0 && std::memcmp(cfg, g.cfg, sizeof cfg) == 0) hit = &g;
        if (!hit) {
            if (b->graphs.size() >= 8) {
                auto old = b->graphs.begin();
                for (auto it = b->graphs.begin(); it != b->graphs.end(); ++it)
                    if (it->used < old->used) old = it;
                HIP_CHECK(hipStreamSynchronize(st));
                HIP_CHECK(hipGraphExecDestroy(old->exec));
                HIP_CHECK(hipGraphDestroy(old->graph));
                b->graphs.erase(old);
            }
            gis_batch::GraphEntry e{};
            e.key = key;
            e.epoch = ctx->ix.epoch;
            std::memcpy(e.cfg, cfg, sizeof cfg);
            HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            try {
                enqueue_run(b, r.mode, r.rescue_mode, r.rcut, r.hard, r.stop);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(st, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIP_CHECK(hipStreamEndCapture(st, &e.graph));
            const hipError_t ie = hipGraphInstantiate(&e.exec, e.graph, nullptr, nullptr, 0);
            if (ie != hipSuccess) {
                (void)hipGraphDestroy(e.graph);
                throw HipError{ie, "hipGraphInstantiate"};
            }
            b->graphs.push_back(e);
            hit = &b->graphs.back();
        }
        hit->used = ++b->graph_clock;
        HIP_CHECK(hipGraphLaunch(hit->exec, st));
    } else {
        enqueue_run(b, r.mode, r.rescue_mode, r.rcut, r.hard, r.stop);
    }
    HIP_CHECK(hipGetLastError());
}

bool batch_ready(const gis_batch *b) { return hipStreamQuery(b->stream) == hipSuccess; }

// wait for the run on the stream; returns the overflow bits
static unsigned finish_once(gis_batch *b, const RunOpts &r) {
    HIP_CHECK(hipStreamSynchronize(b->stream));
    const bool fused = r.mode == GIS_MODE_FUSED && r.stop == STOP_NONE;
    b->ws->n_spill = fused ? b->ws->h_ctl->n_spill : 0;
    b->ws->tot_nams = r.stop == STOP_NONE ? b->ws->h_ctl->tot[2] : 0;
    return b->ws->h_ctl->overflow;
}

void batch_complete(gis_batch *b, const RunOpts &r) {
    unsigned ovf = finish_once(b, r);
    for (int attempt = 1; ovf && attempt < 6; attempt++) {
        // after a hit overflow the merges did not run, and after a merge
        // without its scratch some reads are missing, so the NAM total is
        // not final: grow what overflowed now and size the output next time
        if (ovf & (OVF_HITS | OVF_RHITS | OVF_MERGE)) ovf &= ~(unsigned)OVF_OUT;
        if (ovf & OVF_MERGE) b->ws->merge_scratch = true;
        ensure_pools(b, (ovf & OVF_HITS) ? b->ws->h_ctl->tot[0] + b->ws->h_ctl->tot[0] / 8 + 4096 : 0,
                     (ovf & OVF_RHITS) ? b->ws->h_ctl->tot[1] + b->ws->h_ctl->tot[1] / 8 + 4096 : 0,
                     (ovf & OVF_OUT) ? b->ws->tot_nams + b->ws->tot_nams / 8 + 1024 : 0);
        batch_launch(b, r);
        ovf = finish_once(b, r);
    }
    if (ovf) throw HipError{hipErrorOutOfMemory, "batch arrays still too small after regrowth"};
    b->last_mode = r.mode;
    b->last_graph = r.graph;
    b->ran = r.stop == STOP_NONE;
}

void batch_run(gis_batch *b, const gis_map_options &o, int rescue_mode, int stop) {
    const RunOpts r = run_opts(b, o, rescue_mode, stop);
    batch_launch(b, r);
    batch_complete(b, r);
}

int batch_download(gis_batch *b, gis_nam_out *out) {
    gis_ctx *ctx = b->ctx;
    if (!out) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null output");
    if (!b->ran) return fail(ctx, GIS_ERR_STATE, "batch has not been run");
    if (b->ws->owner != b)
        return fail(ctx, GIS_ERR_STATE, "another batch sharing the workspace ran since this batch's run");
    const uint32_t n = b->n_reads;
    out->n_nams = b->ws->tot_nams;
    if (out->nam_capacity < b->ws->tot_nams || (b->ws->tot_nams && !out->nams))
        return fail(ctx, GIS_ERR_CAPACITY,
                    "NAM capacity " + std::to_string(out->nam_capacity) + " < " + std::to_string(b->ws->tot_nams));
    hipStream_t st = b->stream;
    if (out->nam_offsets)
        HIP_CHECK(hipMemcpyAsync(out->nam_offsets, b->ws->nam_off.p, 8ull * (n + 1), hipMemcpyDeviceToHost, st));
    if (b->ws->tot_nams)
        HIP_CHECK(hipMemcpyAsync(out->nams, b->ws->nams_out.p, sizeof(gis_nam) * b->ws->tot_nams,
                                 hipMemcpyDeviceToHost, st));
    if (out->nonrepetitive_fraction && n)
        HIP_CHECK(hipMemcpyAsync(out->nonrepetitive_fraction, b->ws->frac.p, 4ull * n, hipMemcpyDeviceToHost, st));
    if (out->flags && n) HIP_CHECK(hipMemcpyAsync(out->flags, b->ws->flags.p, n, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    return GIS_OK;
}

gis_batch *scratch_batch(gis_ctx *ctx) {
    if (!ctx->scratch) ctx->scratch = batch_new(ctx, nullptr);
    return ctx->scratch;
}

gis_map_options opts_or_default(const gis_map_options *opt) {
    gis_map_options o;
    if (opt) o = *opt;
    else gis_map_options_default(&o);
    if (o.mode != GIS_MODE_FUSED) o.mode = GIS_MODE_PLAIN;
    return o;
}

}  // namespace gis

// ---------------- stage entry points ----------------

extern "C" int gis_syncmers_batch(gis_ctx *ctx, const char *seqs, const uint64_t *offsets,
                                  const uint32_t *lens, uint32_t n_reads, uint64_t *out_offsets,
                                  gis_syncmer *out, uint64_t capacity, uint64_t *n_out) {
    API_BEGIN(ctx)
    gis_batch *b = scratch_batch(ctx);
    int rc = batch_upload(b, seqs, offsets, lens, n_reads, true);
    if (rc) return rc;
    const gis_map_options o = opts_or_default(nullptr);
    batch_run(b, o, 0, STOP_SEED);
    std::vector<uint32_t> cnt(n_reads);
    if (n_reads) HIP_CHECK(hipMemcpy(cnt.data(), b->ws->sy_cnt.p, 4ull * n_reads, hipMemcpyDeviceToHost));
    uint64_t total = 0;
    for (uint32_t r = 0; r < n_reads; r++) total += cnt[r];
    if (n_out) *n_out = total;
    if (total > capacity || (total && !out)) return fail(ctx, GIS_ERR_CAPACITY, "syncmer capacity too small");
    const uint64_t cap = b->h_sy_off[n_reads];
    std::vector<uint64_t> hh(cap), pp(cap);
    if (cap) {
        HIP_CHECK(hipMemcpy(hh.data(), b->ws->sy_hash.p, 8 * cap, hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(pp.data(), b->ws->sy_pos.p, 8 * cap, hipMemcpyDeviceToHost));
    }
    uint64_t o2 = 0;
    for (uint32_t r = 0; r < n_reads; r++) {
        if (out_offsets) out_offsets[r] = o2;
        for (uint32_t i = 0; i < cnt[r]; i++) {
            out[o2].hash = hh[b->h_sy_off[r] + i];
            out[o2].position = pp[b->h_sy_off[r] + i];
            o2++;
        }
    }
    if (out_offsets) out_offsets[n_reads] = o2;
    return GIS_OK;
    API_END(ctx)
}

// randstrobes_query of every read with the given parameters (the index's
// when params is null); the caller holds the context's lock
static int randstrobes_of(gis_ctx *ctx, const gis_params *params, const char *seqs, const uint64_t *offsets,
                          const uint32_t *lens, uint32_t n_reads, uint64_t *out_offsets, gis_query_randstrobe *out,
                          uint64_t capacity, uint64_t *n_out) {
    gis_batch *b = scratch_batch(ctx);
    struct Reset {
        gis_batch *b;
        ~Reset() { b->seed_params = nullptr; }
    } reset{b};
    b->seed_params = params;
    int rc = batch_upload(b, seqs, offsets, lens, n_reads, false);
    if (rc) return rc;
    const gis_map_options o = opts_or_default(nullptr);
    batch_run(b, o, 0, STOP_SEED);
    std::vector<uint32_t> cnt(n_reads);
    if (n_reads) HIP_CHECK(hipMemcpy(cnt.data(), b->ws->rs_cnt.p, 4ull * n_reads, hipMemcpyDeviceToHost));
    uint64_t total = 0;
    for (uint32_t r = 0; r < n_reads; r++) total += cnt[r];
    if (n_out) *n_out = total;
    if (total > capacity || (total && !out)) return fail(ctx, GIS_ERR_CAPACITY, "randstrobe capacity too small");
    const uint64_t cap = b->tot_rs_cap;
    std::vector<uint64_t> hh(cap);
    std::vector<uint32_t> sp(cap);
    if (cap) {
        HIP_CHECK(hipMemcpy(hh.data(), b->ws->rs_hash.p, 8 * cap, hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(sp.data(), b->ws->rs_span.p, 4 * cap, hipMemcpyDeviceToHost));
    }
    uint64_t o2 = 0;
    for (uint32_t r = 0; r < n_reads; r++) {
        if (out_offsets) out_offsets[r] = o2;
        const uint64_t base = b->h_rs_off[r];
        for (uint32_t i = 0; i < cnt[r]; i++) {
            gis_query_randstrobe q{};
            q.hash = hh[base + i];
            q.start = (int32_t)span_start(sp[base + i]);
            q.end = (int32_t)span_end(sp[base + i]);
            q.is_reverse = i >= cnt[r] / 2;
            out[o2++] = q;
        }
    }
    if (out_offsets) out_offsets[n_reads] = o2;
    return GIS_OK;
}

extern "C" int gis_randstrobes_batch(gis_ctx *ctx, const char *seqs, const uint64_t *offsets,
                                     const uint32_t *lens, uint32_t n_reads, uint64_t *out_offsets,
                                     gis_query_randstrobe *out, uint64_t capacity, uint64_t *n_out) {
    API_BEGIN(ctx)
    return randstrobes_of(ctx, nullptr, seqs, offsets, lens, n_reads, out_offsets, out, capacity, n_out);
    API_END(ctx)
}

extern "C" int gis_randstrobes_params(gis_ctx *ctx, const gis_params *params, const char *seqs,
                                      const uint64_t *offsets, const uint32_t *lens, uint32_t n_reads,
                                      uint64_t *out_offsets, gis_query_randstrobe *out, uint64_t capacity,
                                      uint64_t *n_out) {
    API_BEGIN(ctx)
    if (!params) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null parameters");
    gis_params chk;
    if (gis_params_make(params->canonical_read_length, params->k, params->s, params->l, params->u,
                        (int)params->q, params->max_dist, &chk) != GIS_OK ||
        chk.t_syncmer != params->t_syncmer || chk.w_min != params->w_min || chk.w_max != params->w_max)
        return fail(ctx, GIS_ERR_BAD_PARAMETER, "invalid index parameters");
    return randstrobes_of(ctx, params, seqs, offsets, lens, n_reads, out_offsets, out, capacity, n_out);
    API_END(ctx)
}

extern "C" int gis_lookup_batch(gis_ctx *ctx, const uint64_t *hashes, uint64_t n, uint64_t *positions) {
    API_BEGIN(ctx)
    if (!ctx->ix.loaded) return fail(ctx, GIS_ERR_STATE, "no index loaded");
    if (n && (!hashes || !positions)) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null argument");
    if (n == 0) return GIS_OK;
    gis_batch *b = scratch_batch(ctx);
    uint64_t *dh = ensure<uint64_t>(b->ws->rs_hash, n);
    uint64_t *dp = ensure<uint64_t>(b->ws->pos, n);
    HIP_CHECK(hipMemcpyAsync(dh, hashes, 8 * n, hipMemcpyHostToDevice, ctx->stream));
    launch_find_flat(dev_index(ctx), dh, n, dp, ctx->stream);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(positions, dp, 8 * n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return GIS_OK;
    API_END(ctx)
}

static int copy_hits_out(gis_ctx *ctx, gis_batch *b, uint64_t *out_offsets, gis_hit *out,
                         uint64_t capacity, uint64_t *n_out) {
    const uint32_t n = b->n_reads;
    std::vector<uint8_t> flags(n);
    std::vector<uint64_t> ho(n + 1), rho(n + 1);
    std::vector<uint32_t> hcount(n);
    if (n) HIP_CHECK(hipMemcpy(flags.data(), b->ws->flags.p, n, hipMemcpyDeviceToHost));
    if (n) HIP_CHECK(hipMemcpy(hcount.data(), b->ws->hit_total.p, 4ull * n, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(ho.data(), b->ws->hit_off.p, 8ull * (n + 1), hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(rho.data(), b->ws->rhit_off.p, 8ull * (n + 1), hipMemcpyDeviceToHost));
    std::vector<DevHit> h(ho[n]), rh(rho[n]);
    if (ho[n]) HIP_CHECK(hipMemcpy(h.data(), b->ws->hits.p, sizeof(DevHit) * ho[n], hipMemcpyDeviceToHost));
    if (rho[n]) HIP_CHECK(hipMemcpy(rh.data(), b->ws->rhits.p, sizeof(DevHit) * rho[n], hipMemcpyDeviceToHost));
    uint64_t total = 0;
    // a read's hits: its region or pool place (hit_off) and count (hit_total);
    // rescued reads: the scan-placed rescue array
    for (uint32_t r = 0; r < n; r++)
        total += (flags[r] & GIS_FLAG_RESCUED) ? rho[r + 1] - rho[r] : hcount[r];
    if (n_out) *n_out = total;
    if (total > capacity || (total && !out)) return fail(ctx, GIS_ERR_CAPACITY, "hit capacity too small");
    uint64_t o = 0;
    std::vector<uint64_t> idx;
    for (uint32_t r = 0; r < n; r++) {
        if (out_offsets) out_offsets[r] = o;
        const bool resc = (flags[r] & GIS_FLAG_RESCUED) != 0;
        const std::vector<DevHit> &src = resc ? rh : h;
        const uint64_t lo = resc ? rho[r] : ho[r], hi = resc ? rho[r + 1] : ho[r] + hcount[r];
        // group by (strand, ref) keeping the emission order inside a group
        idx.clear();
        for (uint64_t i = lo; i < hi; i++) idx.push_back(((uint64_t)src[i].ref_strand << 32) | (i - lo));
        std::sort(idx.begin(), idx.end());
        for (uint64_t key : idx) {
            const DevHit &x = src[lo + (uint32_t)key];
            gis_hit g;
            g.query_start = x.query_start;
            g.query_end = x.query_end;
            g.ref_start = x.ref_start;
            g.ref_end = x.ref_end;
            g.ref_id = (int32_t)(x.ref_strand & 0x7fffffffu);
            g.is_rc = (int32_t)(x.ref_strand >> 31);
            out[o++] = g;
        }
    }
    if (out_offsets) out_offsets[n] = o;
    return GIS_OK;
}

extern "C" int gis_hits_batch(gis_ctx *ctx, const char *seqs, const uint64_t *offsets,
                              const uint32_t *lens, uint32_t n_reads, const gis_map_options *opt,
                              uint64_t *out_offsets, gis_hit *out, uint64_t capacity, uint64_t *n_out) {
    API_BEGIN(ctx)
    gis_map_options o = opts_or_default(opt);
    o.mode = GIS_MODE_PLAIN;
    gis_batch *b = scratch_batch(ctx);
    int rc = batch_upload(b, seqs, offsets, lens, n_reads, false);
    if (rc) return rc;
    batch_run(b, o, 0, STOP_HITS);
    return copy_hits_out(ctx, b, out_offsets, out, capacity, n_out);
    API_END(ctx)
}

extern "C" int gis_find_nams_batch(gis_ctx *ctx, const char *seqs, const uint64_t *offsets,
                                   const uint32_t *lens, uint32_t n_reads, const gis_map_options *opt,
                                   gis_nam_out *out) {
    API_BEGIN(ctx)
    const gis_map_options o = opts_or_default(opt);
    gis_batch *b = scratch_batch(ctx);
    int rc = batch_upload(b, seqs, offsets, lens, n_reads, false);
    if (rc) return rc;
    batch_run(b, o, 0, STOP_NONE);
    return batch_download(b, out);
    API_END(ctx)
}

// find_nams / find_nams_rescue over given query randstrobes on the context's
// scratch batch (the caller holds the context's lock)
static int qrs_run(gis_ctx *ctx, const gis_query_randstrobe *qrs, const uint64_t *query_offsets,
                   uint32_t n_queries, int rescue, uint32_t rescue_cutoff, uint32_t hard_cutoff, gis_nam_out *out) {
    if (!ctx->ix.loaded) return fail(ctx, GIS_ERR_STATE, "no index loaded");
    if (!query_offsets || !out) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null argument");
    gis_batch *b = scratch_batch(ctx);
    const uint32_t n = n_queries;
    // each read's queries laid out forward strand first, each strand by (start, end, hash)
    const uint64_t total = query_offsets[n];
    std::vector<uint64_t> hh(total);
    std::vector<uint32_t> sp(total), cnt(n), nf(n);
    std::vector<uint64_t> off(n + 1);
    std::vector<uint64_t> ord;
    for (uint32_t r = 0; r < n; r++) {
        const uint64_t lo = query_offsets[r], hi = query_offsets[r + 1];
        ord.clear();
        for (uint64_t i = lo; i < hi; i++) ord.push_back(i);
        std::stable_sort(ord.begin(), ord.end(), [&](uint64_t a, uint64_t c) {
            const auto &x = qrs[a], &y = qrs[c];
            if (x.is_reverse != y.is_reverse) return x.is_reverse < y.is_reverse;
            if (x.start != y.start) return x.start < y.start;
            if (x.end != y.end) return x.end < y.end;
            return x.hash < y.hash;
        });
        uint32_t f = 0;
        for (uint64_t j = 0; j < ord.size(); j++) {
            const auto &q = qrs[ord[j]];
            if (q.end > GIS_MAX_READ_LEN) return fail(ctx, GIS_ERR_BAD_PARAMETER, "query end too large");
            hh[lo + j] = q.hash;
            sp[lo + j] = span_pack((uint32_t)q.start, (uint32_t)q.end);
            if (!q.is_reverse) f++;
        }
        cnt[r] = (uint32_t)(hi - lo);
        nf[r] = f;
        off[r] = lo;
    }
    off[n] = total;
    b->n_reads = n;
    b->want_sy = false;
    b->any_long = false;
    b->h_rs_off.assign(off.data(), off.size());
    // the per-read randstrobe offsets are the one input array the stages
    // after seeding read (a fresh context's scratch batch has none yet)
    ensure<uint64_t>(b->rs_off, (size_t)n + 1);
    ensure_read_arrays(b, n);
    ensure_slot_arrays(b, total);
    b->ws->owner = b;
    ensure_pools(b, (uint64_t)HIT_REGION * n + 16ull * n + 4096, 16ull * n + 4096, 16ull * n + 1024);
    b->tot_rs_cap = total;
    hipStream_t st = ctx->stream;
    if (total) {
        HIP_CHECK(hipMemcpyAsync(b->ws->rs_hash.p, hh.data(), 8 * total, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(b->ws->rs_span.p, sp.data(), 4 * total, hipMemcpyHostToDevice, st));
    }
    HIP_CHECK(hipMemcpyAsync(b->rs_off.p, off.data(), 8ull * (n + 1), hipMemcpyHostToDevice, st));
    if (n) {
        HIP_CHECK(hipMemcpyAsync(b->ws->rs_cnt.p, cnt.data(), 4ull * n, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(b->ws->rs_nfwd.p, nf.data(), 4ull * n, hipMemcpyHostToDevice, st));
    }
    HIP_CHECK(hipStreamSynchronize(st));
    // the stages after seeding, with the rescue decision forced either way
    const DevIndex ix = dev_index(ctx);
    const ReadSet set{nullptr, nullptr, n};
    uint64_t *tmp = static_cast<uint64_t *>(b->ws->scan_tmp.p);
    for (int attempt = 0; attempt < 6; attempt++) {
        const DevBatch d = make_devbatch(b);
        launch_reset_ctl(d.ctl, st);
        launch_lookup(d, set, ix, rescue ? 2 : 1, st);
        launch_hits(d, set, ix, st);
        launch_rescue_count(d, set, ix, rescue_cutoff, hard_cutoff, st);
        launch_scan(d.rhit_total, d.rhit_off, n, tmp, st);
        launch_rescue_hits(d, set, ix, st);
        launch_merge(d, set, 0, st);
        launch_merge(d, set, 1, st);
        launch_scan(d.nam_cnt, d.nam_off, n, tmp, st);
        launch_compact(d, 0, st);
        launch_ctl_to_host(d.ctl, b->ws->h_ctl_dev, st);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(st));
        b->ws->tot_nams = b->ws->h_ctl->tot[2];
        unsigned ovf = b->ws->h_ctl->overflow;
        if (!ovf) break;
        if (attempt == 5) throw HipError{hipErrorOutOfMemory, "batch arrays still too small after regrowth"};
        if (ovf & (OVF_HITS | OVF_RHITS | OVF_MERGE)) ovf &= ~(unsigned)OVF_OUT;
        if (ovf & OVF_MERGE) b->ws->merge_scratch = true;
        ensure_pools(b, (ovf & OVF_HITS) ? b->ws->h_ctl->tot[0] + 4096 : 0,
                     (ovf & OVF_RHITS) ? b->ws->h_ctl->tot[1] + 4096 : 0,
                     (ovf & OVF_OUT) ? b->ws->tot_nams + 1024 : 0);
    }
    b->ran = true;
    return batch_download(b, out);
}

namespace {

// Concurrent per-read calls (the reference's find_nams is re-entrant and
// called once per read from many threads, step2/nam.cu:553-562) are
// combined: a call queues itself; whichever caller finds the context free
// takes every queued call with the same rescue settings and runs them as one
// batch, then hands each caller its reads' results.
struct QrsCall {
    const gis_query_randstrobe *qrs;
    const uint64_t *qoff;
    uint32_t n;
    int rescue;
    uint32_t rcut, hard;
    gis_nam_out *out;
    int rc = GIS_OK;
    bool done = false;
};

int run_combined(gis_ctx *ctx, std::vector<QrsCall *> &calls) {
    // one batch: the calls' reads back to back
    uint32_t n = 0;
    uint64_t nq = 0;
    for (QrsCall *c : calls) {
        n += c->n;
        nq += c->qoff[c->n] - c->qoff[0];
    }
    std::vector<gis_query_randstrobe> q(nq);
    std::vector<uint64_t> off(n + 1);
    uint64_t at = 0;
    uint32_t r = 0;
    for (QrsCall *c : calls) {
        const uint64_t lo = c->qoff[0], hi = c->qoff[c->n];
        if (hi > lo) std::memcpy(q.data() + at, c->qrs + lo, (hi - lo) * sizeof(gis_query_randstrobe));
        for (uint32_t i = 0; i < c->n; i++) off[r + i] = at + c->qoff[i] - lo;
        at += hi - lo;
        r += c->n;
    }
    off[n] = at;
    std::vector<uint64_t> noff(n + 1);
    std::vector<float> frac(n);
    std::vector<uint8_t> flags(n);
    std::vector<gis_nam> nams(16 + 4ull * at);
    int rc;
    for (;;) {
        gis_nam_out o{noff.data(), nams.data(), nams.size(), frac.data(), flags.data(), 0};
        rc = qrs_run(ctx, q.data(), off.data(), n, calls[0]->rescue, calls[0]->rcut, calls[0]->hard, &o);
        if (rc == GIS_ERR_CAPACITY && o.n_nams > nams.size()) {
            nams.resize(o.n_nams);
            continue;
        }
        break;
    }
    r = 0;
    for (QrsCall *c : calls) {
        if (rc != GIS_OK) {
            c->rc = rc;
            r += c->n;
            continue;
        }
        gis_nam_out *o = c->out;
        const uint64_t a = noff[r], e = noff[r + c->n];
        o->n_nams = e - a;
        if (o->nam_capacity < e - a || (e > a && !o->nams)) {
            c->rc = GIS_ERR_CAPACITY;
        } else {
            if (o->nam_offsets)
                for (uint32_t i = 0; i <= c->n; i++) o->nam_offsets[i] = noff[r + i] - a;
            if (e > a) std::memcpy(o->nams, nams.data() + a, (e - a) * sizeof(gis_nam));
            if (o->nonrepetitive_fraction && c->n) std::memcpy(o->nonrepetitive_fraction, frac.data() + r, 4ull * c->n);
            if (o->flags && c->n) std::memcpy(o->flags, flags.data() + r, c->n);
            c->rc = GIS_OK;
        }
        r += c->n;
    }
    return rc;
}

}  // namespace

extern "C" int gis_find_nams_qrs(gis_ctx *ctx, const gis_query_randstrobe *qrs,
                                 const uint64_t *query_offsets, uint32_t n_queries, int rescue,
                                 uint32_t rescue_cutoff, uint32_t hard_cutoff, gis_nam_out *out) {
    if (!ctx) return GIS_ERR_BAD_PARAMETER;
    if (!query_offsets || !out) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null argument");
    if (n_queries && !qrs && query_offsets[n_queries] > query_offsets[0])
        return fail(ctx, GIS_ERR_BAD_PARAMETER, "null query randstrobes");
    for (uint64_t i = query_offsets[0]; n_queries && i < query_offsets[n_queries]; i++) {
        if (qrs[i].end > GIS_MAX_READ_LEN) return fail(ctx, GIS_ERR_BAD_PARAMETER, "query end too large");
        if (qrs[i].start > qrs[i].end) return fail(ctx, GIS_ERR_BAD_PARAMETER, "query start after its end");
    }
    QrsCall me{qrs, query_offsets, n_queries, rescue ? 1 : 0, rescue_cutoff, hard_cutoff, out};
    std::unique_lock<std::mutex> l(ctx->comb_mu);
    ctx->comb_queue.push_back(&me);
    while (!me.done) {
        if (ctx->comb_busy) {
            ctx->comb_cv.wait(l);
            continue;
        }
        // this caller runs the queued calls that share its rescue settings
        ctx->comb_busy = true;
        std::vector<QrsCall *> take, keep;
        for (void *v : ctx->comb_queue) {
            QrsCall *c = static_cast<QrsCall *>(v);
            const bool same = c->rescue == me.rescue && (!me.rescue || (c->rcut == me.rcut && c->hard == me.hard));
            (same ? take : keep).push_back(c);
        }
        ctx->comb_queue.assign(keep.begin(), keep.end());
        l.unlock();
        int rc = GIS_OK;
        {
            std::lock_guard<std::recursive_mutex> lk(ctx->mu);
            try {
                HIP_CHECK(hipSetDevice(ctx->device));
                rc = run_combined(ctx, take);
            } catch (const HipError &e) {
                rc = hip_fail(ctx, e);
            } catch (const std::bad_alloc &) {
                rc = fail(ctx, GIS_ERR_NOMEM, "host allocation failed");
            }
        }
        l.lock();
        for (QrsCall *c : take) {
            if (rc != GIS_OK) c->rc = rc;
            c->done = true;
        }
        ctx->comb_busy = false;
        ctx->comb_cv.notify_all();
    }
    return me.rc;
}

// ---------------- device-resident batches ----------------

static const char *const IN_FLIGHT_MSG = "a launched run on this batch's workspace was not waited for";

extern "C" int gis_batch_create(gis_ctx *ctx, uint32_t max_reads, uint64_t max_bases, gis_batch **out) {
    API_BEGIN(ctx)
    if (!out) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null output");
    gis_batch *b = batch_new(ctx, nullptr);
    ensure_input_arrays(b, max_reads);
    ensure_read_arrays(b, max_reads);
    ensure<char>(b->seqs, max_bases);
    *out = b;
    return GIS_OK;
    API_END(ctx)
}

extern "C" int gis_batch_create_async(gis_ctx *ctx, uint32_t max_reads, uint64_t max_bases, gis_batch **out) {
    API_BEGIN(ctx)
    if (!out) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null output");
    gis_batch *b = batch_new_stream(ctx);
    try {
        ensure_input_arrays(b, max_reads);
        ensure_read_arrays(b, max_reads);
        ensure<char>(b->seqs, max_bases);
    } catch (...) {
        batch_free(b);
        throw;
    }
    *out = b;
    return GIS_OK;
    API_END(ctx)
}

extern "C" int gis_batch_create_shared(gis_ctx *ctx, uint32_t max_reads, uint64_t max_bases,
                                       gis_batch *share_with, gis_batch **out) {
    API_BEGIN(ctx)
    if (!out || !share_with) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null argument");
    if (share_with->ctx != ctx) return fail(ctx, GIS_ERR_BAD_PARAMETER, "the batch belongs to another context");
    gis_batch *b = batch_new(ctx, share_with->ws);
    b->stream = share_with->stream;
    ensure_input_arrays(b, max_reads);
    ensure<char>(b->seqs, max_bases);
    *out = b;
    return GIS_OK;
    API_END(ctx)
}

extern "C" void gis_batch_destroy(gis_batch *b) {
    if (!b) return;
    gis_ctx *ctx = b->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(b->stream);
    batch_free(b);
}

extern "C" int gis_batch_upload(gis_batch *b, const char *seqs, const uint64_t *offsets,
                                const uint32_t *lens, uint32_t n_reads) {
    if (!b) return GIS_ERR_BAD_PARAMETER;
    gis_ctx *ctx = b->ctx;
    API_BEGIN(ctx)
    if (b->ws->in_flight) return fail(ctx, GIS_ERR_STATE, IN_FLIGHT_MSG);
    return batch_upload(b, seqs, offsets, lens, n_reads, false);
    API_END(ctx)
}

extern "C" int gis_batch_run(gis_batch *b, const gis_map_options *opt) {
    if (!b) return GIS_ERR_BAD_PARAMETER;
    gis_ctx *ctx = b->ctx;
    API_BEGIN(ctx)
    if (!ctx->ix.loaded) return fail(ctx, GIS_ERR_STATE, "no index loaded");
    if (b->ws->in_flight) return fail(ctx, GIS_ERR_STATE, IN_FLIGHT_MSG);
    batch_run(b, opts_or_default(opt), 0, STOP_NONE);
    return GIS_OK;
    API_END(ctx)
}

extern "C" int gis_batch_launch(gis_batch *b, const gis_map_options *opt) {
    if (!b) return GIS_ERR_BAD_PARAMETER;
    gis_ctx *ctx = b->ctx;
    API_BEGIN(ctx)
    if (!ctx->ix.loaded) return fail(ctx, GIS_ERR_STATE, "no index loaded");
    if (b->ws->in_flight) return fail(ctx, GIS_ERR_STATE, IN_FLIGHT_MSG);
    b->pending_opt = opts_or_default(opt);
    batch_launch(b, run_opts(b, b->pending_opt, 0, STOP_NONE));
    b->ws->in_flight = b;
    return GIS_OK;
    API_END(ctx)
}

extern "C" int gis_batch_wait(gis_batch *b) {
    if (!b) return GIS_ERR_BAD_PARAMETER;
    gis_ctx *ctx = b->ctx;
    API_BEGIN(ctx)
    if (b->ws->in_flight != b) return fail(ctx, GIS_ERR_STATE, "no launched run of this batch to wait for");
    // released first: after an error the workspace takes runs again
    b->ws->in_flight = nullptr;
    batch_complete(b, run_opts(b, b->pending_opt, 0, STOP_NONE));
    return GIS_OK;
    API_END(ctx)
}

extern "C" int gis_batch_download(gis_batch *b, gis_nam_out *out) {
    if (!b) return GIS_ERR_BAD_PARAMETER;
    gis_ctx *ctx = b->ctx;
    API_BEGIN(ctx)
    return batch_download(b, out);
    API_END(ctx)
}

extern "C" int gis_batch_nam_count(gis_batch *b, uint64_t *n_nams) {
    if (!b || !n_nams) return GIS_ERR_BAD_PARAMETER;
    if (b->ws->owner != b || b->ws->in_flight) return GIS_ERR_STATE;
    *n_nams = b->ws->tot_nams;
    return GIS_OK;
}

extern "C" int gis_batch_stage_times(gis_batch *b, double *ms, int cap, int *n, char *names, int names_cap) {
    if (!b) return GIS_ERR_BAD_PARAMETER;
    gis_ctx *ctx = b->ctx;
    API_BEGIN(ctx)
    if (b->ws->in_flight) return fail(ctx, GIS_ERR_STATE, IN_FLIGHT_MSG);
    if (b->last_graph)
        return fail(ctx, GIS_ERR_STATE, "stage times need a run with GIS_RUN_NO_GRAPH (graph replays carry no events)");
    std::string nm;
    int k = 0;
    for (int i = 1; i < b->n_ev && k < cap; i++, k++) {
        float t = 0;
        HIP_CHECK(hipEventElapsedTime(&t, b->ev[i - 1], b->ev[i]));
        if (ms) ms[k] = t;
        if (!nm.empty()) nm += ",";
        nm += b->ev_names[i];
    }
    if (n) *n = k;
    if (names && names_cap > 0) std::snprintf(names, (size_t)names_cap, "%s", nm.c_str());
    return GIS_OK;
    API_END(ctx)
}

extern "C" int gis_batch_lookup_bytes(gis_batch *b, uint64_t *bytes, uint64_t *n_lookups) {
    if (!b) return GIS_ERR_BAD_PARAMETER;
    gis_ctx *ctx = b->ctx;
    API_BEGIN(ctx)
    if (b->ws->owner != b) return fail(ctx, GIS_ERR_STATE, "another batch ran on the shared workspace since");
    if (!b->ran) return fail(ctx, GIS_ERR_STATE, "batch has not been run");
    if (b->last_mode != GIS_MODE_PLAIN) return fail(ctx, GIS_ERR_STATE, "lookup bytes are counted by the per-stage path");
    const uint32_t n = b->n_reads;
    std::vector<uint32_t> by(n), cnt(n);
    if (n) {
        HIP_CHECK(hipMemcpy(by.data(), b->ws->lk_bytes.p, 4ull * n, hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(cnt.data(), b->ws->rs_cnt.p, 4ull * n, hipMemcpyDeviceToHost));
    }
    uint64_t tb = 0, tq = 0;
    for (uint32_t r = 0; r < n; r++) {
        tb += by[r];
        tq += cnt[r];
    }
    if (bytes) *bytes = tb;
    if (n_lookups) *n_lookups = tq;
    return GIS_OK;
    API_END(ctx)
}

extern "C" int gis_batch_lookup_streams(gis_batch *b, uint64_t *stream_bytes) {
    if (!b || !stream_bytes) return GIS_ERR_BAD_PARAMETER;
    gis_ctx *ctx = b->ctx;
    API_BEGIN(ctx)
    if (b->ws->owner != b) return fail(ctx, GIS_ERR_STATE, "another batch ran on the shared workspace since");
    if (!b->ran) return fail(ctx, GIS_ERR_STATE, "batch has not been run");
    if (b->last_mode != GIS_MODE_PLAIN) return fail(ctx, GIS_ERR_STATE, "lookup bytes are counted by the per-stage path");
    const uint32_t n = b->n_reads;
    std::vector<uint32_t> sb(n);
    if (n) HIP_CHECK(hipMemcpy(sb.data(), b->ws->lk_stream.p, 4ull * n, hipMemcpyDeviceToHost));
    uint64_t t = 0;
    for (uint32_t r = 0; r < n; r++) t += sb[r];
    *stream_bytes = t;
    return GIS_OK;
    API_END(ctx)
}

// ---------------- autotuner of the fused form ----------------

extern "C" int gis_batch_autotune(gis_batch *b, const gis_map_options *opt, gis_fused_config *best,
                                  double *best_ms, uint32_t *n_tried) {
    if (!b) return GIS_ERR_BAD_PARAMETER;
    gis_ctx *ctx = b->ctx;
    API_BEGIN(ctx)
    if (!ctx->ix.loaded) return fail(ctx, GIS_ERR_STATE, "no index loaded");
    if (b->n_reads == 0) return fail(ctx, GIS_ERR_STATE, "no reads uploaded");
    if (b->ws->in_flight) return fail(ctx, GIS_ERR_STATE, IN_FLIGHT_MSG);
    gis_map_options o = opts_or_default(opt);
    o.mode = GIS_MODE_FUSED;
    o.flags |= GIS_RUN_NO_GRAPH;
    // candidates: waves per block x length split x queue discipline
    struct Q { int queue, chunk, bpc; };
    const Q queues[] = {{0, 0, 0}, {0, 0, 4}, {0, 0, 8}, {1, 4, 4}, {1, 4, 8}, {1, 16, 8}, {1, 64, 4}};
    std::vector<gis_fused_config> cands;
    for (int w : {2, 4, 8})
        for (int split : {0, 1})
            for (const Q &q : queues) cands.push_back(gis_fused_config{w, split, q.queue, q.chunk, q.bpc});
    const gis_fused_config saved = ctx->fcfg;
    double bt = 1e300;
    gis_fused_config bc = saved;
    try {
        for (const gis_fused_config &c : cands) {
            ctx->fcfg = c;
            batch_run(b, o, 0, STOP_NONE);  // first launch of the variant; pools sized
            double t = 1e300;
            for (int rep = 0; rep < 3; rep++) {
                const auto t0 = std::chrono::steady_clock::now();
                batch_run(b, o, 0, STOP_NONE);
                const double ms =
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                t = std::min(t, ms);
            }
            if (t < bt) {
                bt = t;
                bc = c;
            }
        }
    } catch (...) {
        ctx->fcfg = saved;
        throw;
    }
    ctx->fcfg = bc;
    if (best) *best = bc;
    if (best_ms) *best_ms = bt;
    if (n_tried) *n_tried = (uint32_t)cands.size();
    return GIS_OK;
    API_END(ctx)
}
