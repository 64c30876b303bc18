// Host driver and C ABI of the seeding path: contexts, the resident index,
// device batches, the stage entry points and the whole-path NAM entry point.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gis.h"
#include "gis_core.h"
#include "index_host.h"
#include "kernels.h"

using namespace gis;

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct HipError {
    hipError_t e;
    const char *what;
};

#define HIP_CHECK(call)                                             \
    do {                                                            \
        hipError_t e_ = (call);                                     \
        if (e_ != hipSuccess) throw HipError{e_, #call};            \
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
        HIP_CHECK(hipMalloc(&b.p, alloc));
        b.bytes = alloc;
    }
    return static_cast<T *>(b.p);
}

void release(DevBuf &b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

}  // namespace

struct gis_batch;

struct DevIndexStore {
    DevBuf rs, starts;
    uint64_t n = 0;
    int bits = 0;
    uint32_t filter_cutoff = 0;
    gis_params params{};
    uint64_t distinct = 0, once = 0, high = 0, mid = 0, index_cutoff = 0;
    bool loaded = false;
};

struct gis_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::recursive_mutex mu;
    std::string err;
    DevIndexStore ix;
    gis_batch *scratch = nullptr;  // batch reused by the host-buffer entry points
};

struct gis_batch {
    gis_ctx *ctx = nullptr;
    uint32_t max_reads = 0;
    uint64_t max_bases = 0;
    uint32_t n_reads = 0;
    bool want_sy = false;
    bool any_long = false;
    // host staging
    std::vector<uint64_t> h_rs_off, h_sy_off, h_long_off;
    // device buffers
    DevBuf seqs, read_off, read_len, rs_off, sy_off, long_off, long_scratch;
    DevBuf rs_cnt, rs_nfwd, rs_hash, rs_start, rs_end, sy_cnt, sy_hash, sy_pos;
    DevBuf pos, lk_bytes, hcnt, rcnt, hit_total, rhit_total, nam_cnt, frac, flags;
    DevBuf hit_off, rhit_off, nam_off, hits, rhits, keys, rkeys, aux, raux;
    DevBuf nam_ws, nam_fin, rnam_ws, rnam_fin, nams_out, scan_tmp, totals;
    // results of the last run
    uint64_t tot_hits = 0, tot_rhits = 0, tot_nams = 0, tot_rs_cap = 0;
    hipEvent_t ev[16] = {};
    int n_ev = 0;
    std::vector<std::string> ev_names;
    bool ran = false;
};

namespace {

int fail(gis_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int hip_fail(gis_ctx *ctx, const HipError &e) {
    return fail(ctx, GIS_ERR_HIP, std::string(e.what) + ": " + hipGetErrorString(e.e));
}

DevIndex dev_index(const gis_ctx *ctx) {
    DevIndex d;
    d.rs = static_cast<const uint4 *>(ctx->ix.rs.p);
    d.starts = static_cast<const uint64_t *>(ctx->ix.starts.p);
    d.n = ctx->ix.n;
    d.bits = ctx->ix.bits;
    d.filter_cutoff = ctx->ix.filter_cutoff;
    d.k = ctx->ix.params.k;
    return d;
}

// long-read scratch bytes of the seeding stage for one read of length L
uint64_t long_bytes(uint64_t L) {
    auto a8 = [](uint64_t x) { return (x + 7) & ~7ULL; };
    return 16 * L + 8 * (L / 64 + 2) + a8(4 * L) + a8(L) + a8(L) + 64;
}

void upload_index(gis_ctx *ctx, const HostIndex &h) {
    DevIndexStore &ix = ctx->ix;
    ix.loaded = false;
    HIP_CHECK(hipSetDevice(ctx->device));
    // drop the previous table first so two tables never coexist in HBM
    release(ix.rs);
    release(ix.starts);
    ensure<gis_ref_randstrobe>(ix.rs, h.rs.size());
    ensure<uint64_t>(ix.starts, h.starts.size());
    if (!h.rs.empty())
        HIP_CHECK(hipMemcpy(ix.rs.p, h.rs.data(), h.rs.size() * sizeof(gis_ref_randstrobe),
                            hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(ix.starts.p, h.starts.data(), h.starts.size() * 8, hipMemcpyHostToDevice));
    ix.n = h.rs.size();
    ix.bits = h.bits;
    ix.filter_cutoff = h.filter_cutoff;
    ix.params = h.params;
    ix.distinct = h.distinct_strobemers;
    ix.once = h.tot_occur_once;
    ix.high = h.tot_high_ab;
    ix.mid = h.tot_mid_ab;
    ix.index_cutoff = h.index_cutoff;
    ix.loaded = true;
}

DevBatch make_devbatch(gis_batch *b) {
    DevBatch d{};
    d.seqs = static_cast<const char *>(b->seqs.p);
    d.read_off = static_cast<const uint64_t *>(b->read_off.p);
    d.read_len = static_cast<const uint32_t *>(b->read_len.p);
    d.n_reads = b->n_reads;
    d.rs_off = static_cast<const uint64_t *>(b->rs_off.p);
    d.rs_cnt = static_cast<uint32_t *>(b->rs_cnt.p);
    d.rs_nfwd = static_cast<uint32_t *>(b->rs_nfwd.p);
    d.rs_hash = static_cast<uint64_t *>(b->rs_hash.p);
    d.rs_start = static_cast<uint32_t *>(b->rs_start.p);
    d.rs_end = static_cast<uint32_t *>(b->rs_end.p);
    d.sy_off = static_cast<const uint64_t *>(b->sy_off.p);
    d.sy_cnt = static_cast<uint32_t *>(b->sy_cnt.p);
    d.sy_hash = static_cast<uint64_t *>(b->sy_hash.p);
    d.sy_pos = static_cast<uint64_t *>(b->sy_pos.p);
    d.pos = static_cast<uint64_t *>(b->pos.p);
    d.hcnt = static_cast<uint32_t *>(b->hcnt.p);
    d.lk_bytes = static_cast<uint32_t *>(b->lk_bytes.p);
    d.rcnt = static_cast<uint32_t *>(b->rcnt.p);
    d.hit_total = static_cast<uint32_t *>(b->hit_total.p);
    d.frac = static_cast<float *>(b->frac.p);
    d.flags = static_cast<uint8_t *>(b->flags.p);
    d.rhit_total = static_cast<uint32_t *>(b->rhit_total.p);
    d.hit_off = static_cast<const uint64_t *>(b->hit_off.p);
    d.hits = static_cast<DevHit *>(b->hits.p);
    d.rhit_off = static_cast<const uint64_t *>(b->rhit_off.p);
    d.rhits = static_cast<DevHit *>(b->rhits.p);
    d.keys = static_cast<uint64_t *>(b->keys.p);
    d.aux = static_cast<uint32_t *>(b->aux.p);
    d.nam_ws = static_cast<gis_nam *>(b->nam_ws.p);
    d.nam_fin = static_cast<gis_nam *>(b->nam_fin.p);
    d.rnam_ws = static_cast<gis_nam *>(b->rnam_ws.p);
    d.rnam_fin = static_cast<gis_nam *>(b->rnam_fin.p);
    d.rkeys = static_cast<uint64_t *>(b->rkeys.p);
    d.raux = static_cast<uint32_t *>(b->raux.p);
    d.nam_cnt = static_cast<uint32_t *>(b->nam_cnt.p);
    d.nam_off = static_cast<const uint64_t *>(b->nam_off.p);
    d.nams_out = static_cast<gis_nam *>(b->nams_out.p);
    d.long_off = static_cast<const uint64_t *>(b->long_off.p);
    d.long_scratch = static_cast<uint8_t *>(b->long_scratch.p);
    return d;
}

// per-read buffers sized for n reads
void ensure_read_arrays(gis_batch *b, uint32_t n) {
    ensure<uint64_t>(b->read_off, n);
    ensure<uint32_t>(b->read_len, n);
    ensure<uint64_t>(b->rs_off, n + 1);
    ensure<uint64_t>(b->sy_off, n + 1);
    ensure<uint64_t>(b->long_off, n + 1);
    ensure<uint32_t>(b->rs_cnt, n);
    ensure<uint32_t>(b->rs_nfwd, n);
    ensure<uint32_t>(b->sy_cnt, n);
    ensure<uint32_t>(b->hit_total, n);
    ensure<uint32_t>(b->lk_bytes, n);
    ensure<uint32_t>(b->rhit_total, n);
    ensure<uint32_t>(b->nam_cnt, n);
    ensure<float>(b->frac, n);
    ensure<uint8_t>(b->flags, n);
    ensure<uint64_t>(b->hit_off, n + 1);
    ensure<uint64_t>(b->rhit_off, n + 1);
    ensure<uint64_t>(b->nam_off, n + 1);
    ensure<uint64_t>(b->scan_tmp, scan_tmp_words(n) + 8);
    ensure<uint64_t>(b->totals, 8);
}

void ensure_slot_arrays(gis_batch *b, uint64_t slots) {
    ensure<uint64_t>(b->rs_hash, slots);
    ensure<uint32_t>(b->rs_start, slots);
    ensure<uint32_t>(b->rs_end, slots);
    ensure<uint64_t>(b->pos, slots);
    ensure<uint32_t>(b->hcnt, slots);
    ensure<uint32_t>(b->rcnt, slots);
}

void ensure_hit_arrays(gis_batch *b, uint64_t hits, bool rescue) {
    const uint64_t n = b->n_reads;
    if (!rescue) {
        ensure<DevHit>(b->hits, hits);
        ensure<uint64_t>(b->keys, 2 * hits);
        ensure<uint32_t>(b->aux, 2 * hits + n + 1);
        ensure<gis_nam>(b->nam_ws, hits);
        ensure<gis_nam>(b->nam_fin, hits);
    } else {
        ensure<DevHit>(b->rhits, hits);
        ensure<uint64_t>(b->rkeys, 2 * hits);
        ensure<uint32_t>(b->raux, 2 * hits + n + 1);
        ensure<gis_nam>(b->rnam_ws, hits);
        ensure<gis_nam>(b->rnam_fin, hits);
    }
}

uint64_t read_total(gis_batch *b, const DevBuf &offsets, uint32_t n) {
    uint64_t t = 0;
    HIP_CHECK(hipMemcpyAsync(&t, static_cast<const uint64_t *>(offsets.p) + n, 8,
                             hipMemcpyDeviceToHost, b->ctx->stream));
    HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
    return t;
}

void mark(gis_batch *b, const char *name) {
    if (b->n_ev >= 16) return;
    if (!b->ev[b->n_ev]) HIP_CHECK(hipEventCreate(&b->ev[b->n_ev]));
    HIP_CHECK(hipEventRecord(b->ev[b->n_ev], b->ctx->stream));
    b->ev_names.push_back(name);
    b->n_ev++;
}

// Stage reads into the batch: capacities and offsets are computed here from
// the lengths so the device work needs no size round trip for them.
int batch_upload(gis_batch *b, const char *seqs, const uint64_t *offsets, const uint32_t *lens,
                 uint32_t n, bool want_sy) {
    gis_ctx *ctx = b->ctx;
    if (!ctx->ix.loaded) return fail(ctx, GIS_ERR_STATE, "no index loaded");
    if (n > 0 && (!seqs || !offsets || !lens))
        return fail(ctx, GIS_ERR_BAD_PARAMETER, "null read arrays");
    const gis_params &p = ctx->ix.params;
    uint64_t span = 0;
    b->h_rs_off.assign(n + 1, 0);
    b->h_sy_off.assign(n + 1, 0);
    b->h_long_off.assign(n + 1, 0);
    b->any_long = false;
    for (uint32_t r = 0; r < n; r++) {
        const uint64_t L = lens[r];
        if (L > GIS_MAX_READ_LEN)
            return fail(ctx, GIS_ERR_BAD_PARAMETER,
                        "read " + std::to_string(r) + " longer than " + std::to_string(GIS_MAX_READ_LEN));
        span = std::max(span, offsets[r] + L);
        const int64_t nk = (int64_t)L - p.k + 1;
        const int64_t nf = nk - (int64_t)p.w_min;
        const uint64_t cap_sy = nk > 0 ? (uint64_t)nk : 0;
        const uint64_t cap_rs = (nf > 0 && L >= p.w_max) ? 2 * (uint64_t)nf : 0;
        b->h_rs_off[r + 1] = b->h_rs_off[r] + cap_rs;
        b->h_sy_off[r + 1] = b->h_sy_off[r] + cap_sy;
        const bool lg = L > (uint64_t)LMAX_LDS;
        b->any_long |= lg;
        b->h_long_off[r + 1] = b->h_long_off[r] + (lg ? long_bytes(L) : 0);
    }
    // h_long_off[r] is the byte offset of read r's scratch
    b->n_reads = n;
    b->want_sy = want_sy;
    ensure_read_arrays(b, n);
    const uint64_t slots = b->h_rs_off[n];
    b->tot_rs_cap = slots;
    ensure_slot_arrays(b, slots);
    if (want_sy) {
        ensure<uint64_t>(b->sy_hash, b->h_sy_off[n]);
        ensure<uint64_t>(b->sy_pos, b->h_sy_off[n]);
    }
    if (b->any_long) ensure<uint8_t>(b->long_scratch, b->h_long_off[n]);
    ensure<char>(b->seqs, span);
    hipStream_t st = ctx->stream;
    if (span) HIP_CHECK(hipMemcpyAsync(b->seqs.p, seqs, span, hipMemcpyHostToDevice, st));
    if (n) {
        HIP_CHECK(hipMemcpyAsync(b->read_off.p, offsets, 8ull * n, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(b->read_len.p, lens, 4ull * n, hipMemcpyHostToDevice, st));
    }
    HIP_CHECK(hipMemcpyAsync(b->rs_off.p, b->h_rs_off.data(), 8ull * (n + 1), hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(b->sy_off.p, b->h_sy_off.data(), 8ull * (n + 1), hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(b->long_off.p, b->h_long_off.data(), 8ull * (n + 1), hipMemcpyHostToDevice, st));
    HIP_CHECK(hipStreamSynchronize(st));
    b->ran = false;
    return GIS_OK;
}

uint32_t rescue_cutoff_of(const gis_ctx *ctx, const gis_map_options &o) {
    // step2/main.cu:4204: R < 100 ? R * filter_cutoff : 100
    return o.rescue_level < 100 ? (uint32_t)(o.rescue_level * (int)ctx->ix.filter_cutoff) : 100u;
}

// The per-stage path: one launch per stage, sizes of the variable-length
// stages read back between launches.
void batch_run_plain(gis_batch *b, const gis_map_options &o, int rescue_mode, int stop_after) {
    gis_ctx *ctx = b->ctx;
    hipStream_t st = ctx->stream;
    const uint32_t n = b->n_reads;
    const DevIndex ix = dev_index(ctx);
    const SeedParams sp = make_seed_params(ctx->ix.params);
    const uint32_t rcut = rescue_cutoff_of(ctx, o);
    b->n_ev = 0;
    b->ev_names.clear();
    DevBatch d = make_devbatch(b);
    mark(b, "start");
    launch_seed(d, sp, b->want_sy ? 1 : 0, b->any_long ? 1 : 0, st);
    mark(b, "seed");
    if (stop_after == 0) {
        HIP_CHECK(hipStreamSynchronize(st));
        return;
    }
    launch_lookup(d, ix, rescue_mode, st);
    mark(b, "lookup");
    launch_scan(d.hit_total, static_cast<uint64_t *>(b->hit_off.p), n,
                static_cast<uint64_t *>(b->scan_tmp.p), st);
    b->tot_hits = read_total(b, b->hit_off, n);
    ensure_hit_arrays(b, b->tot_hits, false);
    d = make_devbatch(b);
    mark(b, "scan_hits");
    launch_hits(d, ix, st);
    mark(b, "hits");
    launch_rescue_count(d, ix, rcut, o.hard_cutoff, st);
    mark(b, "rescue_count");
    launch_scan(d.rhit_total, static_cast<uint64_t *>(b->rhit_off.p), n,
                static_cast<uint64_t *>(b->scan_tmp.p), st);
    b->tot_rhits = read_total(b, b->rhit_off, n);
    ensure_hit_arrays(b, b->tot_rhits, true);
    d = make_devbatch(b);
    mark(b, "scan_rescue");
    launch_rescue_hits(d, ix, rcut, o.hard_cutoff, st);
    mark(b, "rescue_hits");
    if (stop_after == 1) {
        HIP_CHECK(hipStreamSynchronize(st));
        return;
    }
    launch_merge(d, 0, ix.k, st);
    mark(b, "merge");
    launch_merge(d, 1, ix.k, st);
    mark(b, "rescue_merge");
    launch_scan(d.nam_cnt, static_cast<uint64_t *>(b->nam_off.p), n,
                static_cast<uint64_t *>(b->scan_tmp.p), st);
    b->tot_nams = read_total(b, b->nam_off, n);
    ensure<gis_nam>(b->nams_out, b->tot_nams);
    d = make_devbatch(b);
    mark(b, "scan_nams");
    launch_compact(d, st);
    mark(b, "compact");
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(st));
    b->ran = true;
}

int batch_download(gis_batch *b, gis_nam_out *out) {
    gis_ctx *ctx = b->ctx;
    if (!out) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null output");
    if (!b->ran) return fail(ctx, GIS_ERR_STATE, "batch has not been run");
    const uint32_t n = b->n_reads;
    out->n_nams = b->tot_nams;
    if (out->nam_capacity < b->tot_nams || (b->tot_nams && !out->nams))
        return fail(ctx, GIS_ERR_CAPACITY,
                    "NAM capacity " + std::to_string(out->nam_capacity) + " < " + std::to_string(b->tot_nams));
    hipStream_t st = ctx->stream;
    if (out->nam_offsets)
        HIP_CHECK(hipMemcpyAsync(out->nam_offsets, b->nam_off.p, 8ull * (n + 1), hipMemcpyDeviceToHost, st));
    if (b->tot_nams)
        HIP_CHECK(hipMemcpyAsync(out->nams, b->nams_out.p, sizeof(gis_nam) * b->tot_nams,
                                 hipMemcpyDeviceToHost, st));
    if (out->nonrepetitive_fraction && n)
        HIP_CHECK(hipMemcpyAsync(out->nonrepetitive_fraction, b->frac.p, 4ull * n, hipMemcpyDeviceToHost, st));
    if (out->flags && n) HIP_CHECK(hipMemcpyAsync(out->flags, b->flags.p, n, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    return GIS_OK;
}

gis_batch *new_batch(gis_ctx *ctx) {
    gis_batch *b = new gis_batch();
    b->ctx = ctx;
    return b;
}

void free_batch(gis_batch *b) {
    if (!b) return;
    DevBuf *all[] = {&b->seqs, &b->read_off, &b->read_len, &b->rs_off, &b->sy_off, &b->long_off,
                     &b->long_scratch, &b->rs_cnt, &b->rs_nfwd, &b->rs_hash, &b->rs_start,
                     &b->rs_end, &b->sy_cnt, &b->sy_hash, &b->sy_pos, &b->pos, &b->lk_bytes, &b->hcnt, &b->rcnt,
                     &b->hit_total, &b->rhit_total, &b->nam_cnt, &b->frac, &b->flags, &b->hit_off,
                     &b->rhit_off, &b->nam_off, &b->hits, &b->rhits, &b->keys, &b->rkeys, &b->aux,
                     &b->raux, &b->nam_ws, &b->nam_fin, &b->rnam_ws, &b->rnam_fin, &b->nams_out,
                     &b->scan_tmp, &b->totals};
    for (DevBuf *x : all) release(*x);
    for (auto &e : b->ev)
        if (e) (void)hipEventDestroy(e);
    delete b;
}

gis_batch *scratch_batch(gis_ctx *ctx) {
    if (!ctx->scratch) ctx->scratch = new_batch(ctx);
    return ctx->scratch;
}

#define API_BEGIN(ctx)                                              \
    if (!(ctx)) return GIS_ERR_BAD_PARAMETER;                       \
    std::lock_guard<std::recursive_mutex> lock_((ctx)->mu);         \
    try {                                                           \
        HIP_CHECK(hipSetDevice((ctx)->device));
#define API_END(ctx)                                                \
    }                                                               \
    catch (const HipError &e) { return hip_fail((ctx), e); }        \
    catch (const std::bad_alloc &) { return fail((ctx), GIS_ERR_NOMEM, "host allocation failed"); }

}  // namespace

// ============================================================================
// C ABI

extern "C" int gis_device_count(int *count) {
    if (!count) return GIS_ERR_BAD_PARAMETER;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    return GIS_OK;
}

extern "C" int gis_ctx_create(int device, gis_ctx **out) {
    if (!out) return GIS_ERR_BAD_PARAMETER;
    *out = nullptr;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess || c <= 0) return GIS_ERR_HIP;
    if (device < 0 || device >= c) return GIS_ERR_BAD_PARAMETER;
    if (hipSetDevice(device) != hipSuccess) return GIS_ERR_HIP;
    gis_ctx *ctx = new gis_ctx();
    ctx->device = device;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return GIS_ERR_HIP;
    }
    *out = ctx;
    return GIS_OK;
}

extern "C" void gis_ctx_destroy(gis_ctx *ctx) {
    if (!ctx) return;
    {
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        free_batch(ctx->scratch);
        ctx->scratch = nullptr;
        release(ctx->ix.rs);
        release(ctx->ix.starts);
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}

extern "C" const char *gis_last_error(const gis_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" int gis_index_build(gis_ctx *ctx, const char *const *seqs, const uint64_t *lens,
                               uint32_t n_refs, const gis_params *params, int bits, float f,
                               int n_threads) {
    API_BEGIN(ctx)
    if (!params || (n_refs && (!seqs || !lens))) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null argument");
    uint64_t total = 0;
    for (uint32_t j = 0; j < n_refs; j++) total += lens[j];
    if (total == 0) return fail(ctx, GIS_ERR_INVALID_FASTA, "No reference sequences found");
    HostIndex h;
    std::string err;
    int rc = build_host_index(seqs, lens, n_refs, *params, bits, f, n_threads, h, err);
    if (rc) return fail(ctx, rc, err);
    upload_index(ctx, h);
    return GIS_OK;
    API_END(ctx)
}

extern "C" int gis_index_upload(gis_ctx *ctx, const gis_ref_randstrobe *randstrobes, uint64_t n,
                                const uint64_t *starts, int bits, uint32_t filter_cutoff,
                                const gis_params *params) {
    API_BEGIN(ctx)
    if (!params || !starts || (n && !randstrobes)) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null argument");
    if (bits < 8 || bits > 31) return fail(ctx, GIS_ERR_BAD_PARAMETER, "Bits must be between 8 and 31");
    // the lookups and the hit order rely on the table's (hash, position) order
    for (uint64_t i = 1; i < n; i++) {
        const auto &a = randstrobes[i - 1], &c = randstrobes[i];
        if (a.hash > c.hash || (a.hash == c.hash && a.position > c.position))
            return fail(ctx, GIS_ERR_INVALID_INDEX, "randstrobes not sorted by (hash, position)");
    }
    const uint64_t nb = (1ULL << bits) + 1;
    for (uint64_t i = 0; i < nb; i++)
        if (starts[i] > n || (i && starts[i] < starts[i - 1]))
            return fail(ctx, GIS_ERR_INVALID_INDEX, "bucket starts out of range");
    HostIndex h;
    h.rs.assign(randstrobes, randstrobes + n);
    h.starts.assign(starts, starts + nb);
    h.bits = bits;
    h.filter_cutoff = filter_cutoff;
    h.params = *params;
    upload_index(ctx, h);
    return GIS_OK;
    API_END(ctx)
}

extern "C" int gis_index_read_sti(gis_ctx *ctx, const char *path, const gis_params *expected) {
    API_BEGIN(ctx)
    if (!path) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null path");
    HostIndex h;
    std::string err;
    int rc = read_sti(path, expected, h, err);
    if (rc) return fail(ctx, rc, err);
    upload_index(ctx, h);
    return GIS_OK;
    API_END(ctx)
}

static void download_index(gis_ctx *ctx, HostIndex &h) {
    const DevIndexStore &ix = ctx->ix;
    h.rs.resize(ix.n);
    h.starts.resize((1ULL << ix.bits) + 1);
    if (ix.n)
        HIP_CHECK(hipMemcpy(h.rs.data(), ix.rs.p, ix.n * sizeof(gis_ref_randstrobe), hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(h.starts.data(), ix.starts.p, h.starts.size() * 8, hipMemcpyDeviceToHost));
    h.bits = ix.bits;
    h.filter_cutoff = ix.filter_cutoff;
    h.params = ix.params;
}

extern "C" int gis_index_write_sti(gis_ctx *ctx, const char *path) {
    API_BEGIN(ctx)
    if (!ctx->ix.loaded) return fail(ctx, GIS_ERR_STATE, "no index loaded");
    HostIndex h;
    download_index(ctx, h);
    std::string err;
    int rc = write_sti(path, h, err);
    if (rc) return fail(ctx, rc, err);
    return GIS_OK;
    API_END(ctx)
}

extern "C" int gis_index_info_get(gis_ctx *ctx, gis_index_info *out) {
    API_BEGIN(ctx)
    if (!out) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null output");
    if (!ctx->ix.loaded) return fail(ctx, GIS_ERR_STATE, "no index loaded");
    const DevIndexStore &ix = ctx->ix;
    out->n_randstrobes = ix.n;
    out->bits = ix.bits;
    out->filter_cutoff = ix.filter_cutoff;
    out->params = ix.params;
    out->distinct_strobemers = ix.distinct;
    out->tot_occur_once = ix.once;
    out->tot_high_ab = ix.high;
    out->tot_mid_ab = ix.mid;
    out->index_cutoff = ix.index_cutoff;
    out->device_bytes = ix.rs.bytes + ix.starts.bytes;
    return GIS_OK;
    API_END(ctx)
}

extern "C" int gis_index_download(gis_ctx *ctx, gis_ref_randstrobe *randstrobes, uint64_t *starts) {
    API_BEGIN(ctx)
    if (!ctx->ix.loaded) return fail(ctx, GIS_ERR_STATE, "no index loaded");
    const DevIndexStore &ix = ctx->ix;
    if (randstrobes && ix.n)
        HIP_CHECK(hipMemcpy(randstrobes, ix.rs.p, ix.n * sizeof(gis_ref_randstrobe), hipMemcpyDeviceToHost));
    if (starts)
        HIP_CHECK(hipMemcpy(starts, ix.starts.p, ((1ULL << ix.bits) + 1) * 8, hipMemcpyDeviceToHost));
    return GIS_OK;
    API_END(ctx)
}

extern "C" int gis_syncmers_batch(gis_ctx *ctx, const char *seqs, const uint64_t *offsets,
                                  const uint32_t *lens, uint32_t n_reads, uint64_t *out_offsets,
                                  gis_syncmer *out, uint64_t capacity, uint64_t *n_out) {
    API_BEGIN(ctx)
    gis_batch *b = scratch_batch(ctx);
    int rc = batch_upload(b, seqs, offsets, lens, n_reads, true);
    if (rc) return rc;
    gis_map_options o;
    gis_map_options_default(&o);
    batch_run_plain(b, o, 0, 0);
    std::vector<uint32_t> cnt(n_reads);
    if (n_reads) HIP_CHECK(hipMemcpy(cnt.data(), b->sy_cnt.p, 4ull * n_reads, hipMemcpyDeviceToHost));
    uint64_t total = 0;
    for (uint32_t r = 0; r < n_reads; r++) total += cnt[r];
    if (n_out) *n_out = total;
    if (total > capacity || (total && !out))
        return fail(ctx, GIS_ERR_CAPACITY, "syncmer capacity too small");
    const uint64_t cap = b->h_sy_off[n_reads];
    std::vector<uint64_t> hh(cap), pp(cap);
    if (cap) {
        HIP_CHECK(hipMemcpy(hh.data(), b->sy_hash.p, 8 * cap, hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(pp.data(), b->sy_pos.p, 8 * cap, hipMemcpyDeviceToHost));
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

extern "C" int gis_randstrobes_batch(gis_ctx *ctx, const char *seqs, const uint64_t *offsets,
                                     const uint32_t *lens, uint32_t n_reads, uint64_t *out_offsets,
                                     gis_query_randstrobe *out, uint64_t capacity, uint64_t *n_out) {
    API_BEGIN(ctx)
    gis_batch *b = scratch_batch(ctx);
    int rc = batch_upload(b, seqs, offsets, lens, n_reads, false);
    if (rc) return rc;
    gis_map_options o;
    gis_map_options_default(&o);
    batch_run_plain(b, o, 0, 0);
    std::vector<uint32_t> cnt(n_reads);
    if (n_reads) HIP_CHECK(hipMemcpy(cnt.data(), b->rs_cnt.p, 4ull * n_reads, hipMemcpyDeviceToHost));
    uint64_t total = 0;
    for (uint32_t r = 0; r < n_reads; r++) total += cnt[r];
    if (n_out) *n_out = total;
    if (total > capacity || (total && !out))
        return fail(ctx, GIS_ERR_CAPACITY, "randstrobe capacity too small");
    const uint64_t cap = b->tot_rs_cap;
    std::vector<uint64_t> hh(cap);
    std::vector<uint32_t> ss(cap), ee(cap);
    if (cap) {
        HIP_CHECK(hipMemcpy(hh.data(), b->rs_hash.p, 8 * cap, hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(ss.data(), b->rs_start.p, 4 * cap, hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(ee.data(), b->rs_end.p, 4 * cap, hipMemcpyDeviceToHost));
    }
    uint64_t o2 = 0;
    for (uint32_t r = 0; r < n_reads; r++) {
        if (out_offsets) out_offsets[r] = o2;
        const uint64_t base = b->h_rs_off[r];
        for (uint32_t i = 0; i < cnt[r]; i++) {
            gis_query_randstrobe q{};
            q.hash = hh[base + i];
            q.start = ss[base + i];
            q.end = ee[base + i];
            q.is_reverse = i >= cnt[r] / 2;
            out[o2++] = q;
        }
    }
    if (out_offsets) out_offsets[n_reads] = o2;
    return GIS_OK;
    API_END(ctx)
}

extern "C" int gis_lookup_batch(gis_ctx *ctx, const uint64_t *hashes, uint64_t n, uint64_t *positions) {
    API_BEGIN(ctx)
    if (!ctx->ix.loaded) return fail(ctx, GIS_ERR_STATE, "no index loaded");
    if (n && (!hashes || !positions)) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null argument");
    if (n == 0) return GIS_OK;
    gis_batch *b = scratch_batch(ctx);
    uint64_t *dh = ensure<uint64_t>(b->rs_hash, n);
    uint64_t *dp = ensure<uint64_t>(b->pos, n);
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
    if (n) HIP_CHECK(hipMemcpy(flags.data(), b->flags.p, n, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(ho.data(), b->hit_off.p, 8ull * (n + 1), hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(rho.data(), b->rhit_off.p, 8ull * (n + 1), hipMemcpyDeviceToHost));
    std::vector<DevHit> h(b->tot_hits), rh(b->tot_rhits);
    if (b->tot_hits) HIP_CHECK(hipMemcpy(h.data(), b->hits.p, sizeof(DevHit) * b->tot_hits, hipMemcpyDeviceToHost));
    if (b->tot_rhits)
        HIP_CHECK(hipMemcpy(rh.data(), b->rhits.p, sizeof(DevHit) * b->tot_rhits, hipMemcpyDeviceToHost));
    uint64_t total = 0;
    for (uint32_t r = 0; r < n; r++)
        total += (flags[r] & GIS_FLAG_RESCUED) ? rho[r + 1] - rho[r] : ho[r + 1] - ho[r];
    if (n_out) *n_out = total;
    if (total > capacity || (total && !out)) return fail(ctx, GIS_ERR_CAPACITY, "hit capacity too small");
    uint64_t o = 0;
    for (uint32_t r = 0; r < n; r++) {
        if (out_offsets) out_offsets[r] = o;
        const bool resc = (flags[r] & GIS_FLAG_RESCUED) != 0;
        const std::vector<DevHit> &src = resc ? rh : h;
        const uint64_t lo = resc ? rho[r] : ho[r], hi = resc ? rho[r + 1] : ho[r + 1];
        // group by (strand, ref) keeping the emission order inside a group
        std::vector<uint64_t> idx;
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
    gis_map_options o;
    if (opt) o = *opt;
    else gis_map_options_default(&o);
    gis_batch *b = scratch_batch(ctx);
    int rc = batch_upload(b, seqs, offsets, lens, n_reads, false);
    if (rc) return rc;
    batch_run_plain(b, o, 0, 1);
    return copy_hits_out(ctx, b, out_offsets, out, capacity, n_out);
    API_END(ctx)
}

extern "C" int gis_find_nams_batch(gis_ctx *ctx, const char *seqs, const uint64_t *offsets,
                                   const uint32_t *lens, uint32_t n_reads, const gis_map_options *opt,
                                   gis_nam_out *out) {
    API_BEGIN(ctx)
    gis_map_options o;
    if (opt) o = *opt;
    else gis_map_options_default(&o);
    gis_batch *b = scratch_batch(ctx);
    int rc = batch_upload(b, seqs, offsets, lens, n_reads, false);
    if (rc) return rc;
    batch_run_plain(b, o, 0, 2);
    return batch_download(b, out);
    API_END(ctx)
}

extern "C" int gis_find_nams_qrs(gis_ctx *ctx, const gis_query_randstrobe *qrs,
                                 const uint64_t *query_offsets, uint32_t n_queries, int rescue,
                                 uint32_t rescue_cutoff, uint32_t hard_cutoff, gis_nam_out *out) {
    API_BEGIN(ctx)
    if (!ctx->ix.loaded) return fail(ctx, GIS_ERR_STATE, "no index loaded");
    if (!query_offsets || !out) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null argument");
    gis_batch *b = scratch_batch(ctx);
    const uint32_t n = n_queries;
    // lay the queries out forward strand first, each strand by (start, end, hash)
    const uint64_t total = query_offsets[n];
    std::vector<uint64_t> hh(total);
    std::vector<uint32_t> ss(total), ee(total), cnt(n), nf(n);
    std::vector<uint64_t> off(n + 1);
    for (uint32_t r = 0; r < n; r++) {
        const uint64_t lo = query_offsets[r], hi = query_offsets[r + 1];
        std::vector<uint64_t> ord;
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
            hh[lo + j] = q.hash;
            ss[lo + j] = q.start;
            ee[lo + j] = q.end;
            if (!q.is_reverse) f++;
            if (q.end > GIS_MAX_READ_LEN) return fail(ctx, GIS_ERR_BAD_PARAMETER, "query end too large");
        }
        cnt[r] = (uint32_t)(hi - lo);
        nf[r] = f;
        off[r] = lo;
    }
    off[n] = total;
    b->n_reads = n;
    b->want_sy = false;
    b->any_long = false;
    b->h_rs_off = off;
    ensure_read_arrays(b, n);
    ensure_slot_arrays(b, total);
    b->tot_rs_cap = total;
    hipStream_t st = ctx->stream;
    if (total) {
        HIP_CHECK(hipMemcpyAsync(b->rs_hash.p, hh.data(), 8 * total, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(b->rs_start.p, ss.data(), 4 * total, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(b->rs_end.p, ee.data(), 4 * total, hipMemcpyHostToDevice, st));
    }
    HIP_CHECK(hipMemcpyAsync(b->rs_off.p, off.data(), 8ull * (n + 1), hipMemcpyHostToDevice, st));
    if (n) {
        HIP_CHECK(hipMemcpyAsync(b->rs_cnt.p, cnt.data(), 4ull * n, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(b->rs_nfwd.p, nf.data(), 4ull * n, hipMemcpyHostToDevice, st));
    }
    HIP_CHECK(hipStreamSynchronize(st));
    // rescue_level R gives rescue_cutoff R * filter_cutoff; pass the cutoff
    // directly through a level that reproduces it
    gis_map_options o;
    gis_map_options_default(&o);
    o.hard_cutoff = hard_cutoff;
    // run the stages after seeding
    const DevIndex ix = dev_index(ctx);
    DevBatch d = make_devbatch(b);
    b->n_ev = 0;
    b->ev_names.clear();
    launch_lookup(d, ix, rescue ? 2 : 1, st);
    launch_scan(d.hit_total, static_cast<uint64_t *>(b->hit_off.p), n, static_cast<uint64_t *>(b->scan_tmp.p), st);
    b->tot_hits = read_total(b, b->hit_off, n);
    ensure_hit_arrays(b, b->tot_hits, false);
    d = make_devbatch(b);
    launch_hits(d, ix, st);
    launch_rescue_count(d, ix, rescue_cutoff, hard_cutoff, st);
    launch_scan(d.rhit_total, static_cast<uint64_t *>(b->rhit_off.p), n, static_cast<uint64_t *>(b->scan_tmp.p), st);
    b->tot_rhits = read_total(b, b->rhit_off, n);
    ensure_hit_arrays(b, b->tot_rhits, true);
    d = make_devbatch(b);
    launch_rescue_hits(d, ix, rescue_cutoff, hard_cutoff, st);
    launch_merge(d, 0, ix.k, st);
    launch_merge(d, 1, ix.k, st);
    launch_scan(d.nam_cnt, static_cast<uint64_t *>(b->nam_off.p), n, static_cast<uint64_t *>(b->scan_tmp.p), st);
    b->tot_nams = read_total(b, b->nam_off, n);
    ensure<gis_nam>(b->nams_out, b->tot_nams);
    d = make_devbatch(b);
    launch_compact(d, st);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(st));
    b->ran = true;
    (void)o;
    return batch_download(b, out);
    API_END(ctx)
}

// ---------------- device-resident batches ----------------

extern "C" int gis_batch_create(gis_ctx *ctx, uint32_t max_reads, uint64_t max_bases, gis_batch **out) {
    API_BEGIN(ctx)
    if (!out) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null output");
    gis_batch *b = new_batch(ctx);
    b->max_reads = max_reads;
    b->max_bases = max_bases;
    ensure_read_arrays(b, max_reads);
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
    (void)hipStreamSynchronize(ctx->stream);
    free_batch(b);
}

extern "C" int gis_batch_upload(gis_batch *b, const char *seqs, const uint64_t *offsets,
                                const uint32_t *lens, uint32_t n_reads) {
    if (!b) return GIS_ERR_BAD_PARAMETER;
    gis_ctx *ctx = b->ctx;
    API_BEGIN(ctx)
    return batch_upload(b, seqs, offsets, lens, n_reads, false);
    API_END(ctx)
}

extern "C" int gis_batch_run(gis_batch *b, const gis_map_options *opt) {
    if (!b) return GIS_ERR_BAD_PARAMETER;
    gis_ctx *ctx = b->ctx;
    API_BEGIN(ctx)
    if (!ctx->ix.loaded) return fail(ctx, GIS_ERR_STATE, "no index loaded");
    gis_map_options o;
    if (opt) o = *opt;
    else gis_map_options_default(&o);
    batch_run_plain(b, o, 0, 2);
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
    *n_nams = b->tot_nams;
    return GIS_OK;
}

extern "C" int gis_batch_stage_times(gis_batch *b, double *ms, int cap, int *n, char *names, int names_cap) {
    if (!b) return GIS_ERR_BAD_PARAMETER;
    gis_ctx *ctx = b->ctx;
    API_BEGIN(ctx)
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
    if (names && names_cap > 0) {
        std::snprintf(names, (size_t)names_cap, "%s", nm.c_str());
    }
    return GIS_OK;
    API_END(ctx)
}

extern "C" int gis_batch_lookup_bytes(gis_batch *b, uint64_t *bytes, uint64_t *n_lookups) {
    if (!b) return GIS_ERR_BAD_PARAMETER;
    gis_ctx *ctx = b->ctx;
    API_BEGIN(ctx)
    if (!b->ran) return fail(ctx, GIS_ERR_STATE, "batch has not been run");
    const uint32_t n = b->n_reads;
    std::vector<uint32_t> by(n), cnt(n);
    if (n) {
        HIP_CHECK(hipMemcpy(by.data(), b->lk_bytes.p, 4ull * n, hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(cnt.data(), b->rs_cnt.p, 4ull * n, hipMemcpyDeviceToHost));
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
