// C ABI: contexts and the resident strobemer index (StrobemerIndex,
// step2/index.hpp:37-183).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "driver_internal.h"
#include "index_build.h"
#include "index_host.h"

using namespace gis;

namespace {

// the bucket lines of the resident table (one 128-byte line per bucket)
void build_lines(gis_ctx *ctx) {
    DevIndexStore &ix = ctx->ix;
    ensure<uint4>(ix.lines, (size_t)BUCKET_LINE << ix.bits);
    launch_bucket_lines(static_cast<const uint4 *>(ix.rs.p), static_cast<const uint64_t *>(ix.starts.p), ix.bits,
                        static_cast<uint4 *>(ix.lines.p), ctx->stream);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

void upload_index(gis_ctx *ctx, const HostIndex &h) {
    DevIndexStore &ix = ctx->ix;
    // drop the previous table first so two tables never coexist in HBM
    drop_table(ix);
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
    build_lines(ctx);
    ix.loaded = true;
}

void download_index(gis_ctx *ctx, HostIndex &h) {
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

}  // namespace

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
    // the allocation-free lookup's array address on this device; if it
    // cannot be had, that entry point reports it and the rest works on
    (void)noalloc_setup(device);
    gis_ctx *ctx = new gis_ctx();
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        ctx->n_cu = prop.multiProcessorCount;
    // default-priority stream
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
        batch_free(ctx->scratch);
        ctx->scratch = nullptr;
        for (gis_batch *b : ctx->lanes) batch_free(b);
        ctx->lanes.clear();
        dma_release(ctx->dma);
        ctx->dma = nullptr;
        pairwork_free(ctx->pairs);
        ctx->pairs = nullptr;
        release(ctx->ix.rs);
        release(ctx->ix.starts);
        release(ctx->ix.lines);
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}

extern "C" const char *gis_last_error(const gis_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

// contigs the 24-bit reference index and 32-bit positions of a record can hold
static int check_contigs(gis_ctx *ctx, const char *const *seqs, const uint64_t *lens, uint32_t n_refs) {
    if (n_refs && (!seqs || !lens)) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null argument");
    if (n_refs >= (1u << 24)) return fail(ctx, GIS_ERR_BAD_PARAMETER, "too many reference sequences for the 24-bit ref index");
    for (uint32_t j = 0; j < n_refs; j++) {
        if (lens[j] >= (1ull << 32)) return fail(ctx, GIS_ERR_BAD_PARAMETER, "reference sequence longer than 2^32 - 1");
        if (lens[j] && !seqs[j]) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null reference sequence");
    }
    return GIS_OK;
}

extern "C" int gis_index_build(gis_ctx *ctx, const char *const *seqs, const uint64_t *lens,
                               uint32_t n_refs, const gis_params *params, int bits, float f,
                               int n_threads) {
    (void)n_threads;  // the device seeds every contig
    API_BEGIN(ctx)
    if (!params) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null argument");
    if (int rc = check_contigs(ctx, seqs, lens, n_refs)) return rc;
    uint64_t total = 0;
    for (uint32_t j = 0; j < n_refs; j++) total += lens[j];
    if (total == 0) return fail(ctx, GIS_ERR_INVALID_FASTA, "No reference sequences found");
    if (bits == -1) bits = pick_bits(total, *params);
    if (bits < 8 || bits > 31) return fail(ctx, GIS_ERR_BAD_PARAMETER, "Bits must be between 8 and 31");
    DevIndexStore &ix = ctx->ix;
    drop_table(ix);
    uint64_t n = 0;
    uint64_t hist[RUN_BINS] = {};
    try {
        uint4 *recs = nullptr;
        device_seed_reference(seqs, lens, n_refs, *params, &recs, &n, ctx->stream);
        ix.rs.p = recs;
        ix.rs.bytes = std::max<uint64_t>(n, 1) * sizeof(gis_ref_randstrobe);
        ensure<uint64_t>(ix.starts, (1ull << bits) + 1);
        device_sort_records(static_cast<uint4 *>(ix.rs.p), n, ctx->stream);
        device_finish_index(static_cast<const uint4 *>(ix.rs.p), n, bits, static_cast<uint64_t *>(ix.starts.p),
                            hist, ctx->stream);
    } catch (const BuildError &e) {
        throw HipError{e.e, e.what};
    }
    HostIndex st;
    filter_stats(hist, f, st);
    ix.n = n;
    ix.bits = bits;
    ix.filter_cutoff = st.filter_cutoff;
    ix.params = *params;
    ix.distinct = st.distinct_strobemers;
    ix.once = st.tot_occur_once;
    ix.high = st.tot_high_ab;
    ix.mid = st.tot_mid_ab;
    ix.index_cutoff = st.index_cutoff;
    build_lines(ctx);
    ix.loaded = true;
    return GIS_OK;
    API_END(ctx)
}

extern "C" int gis_reference_randstrobes(gis_ctx *ctx, const char *const *seqs, const uint64_t *lens,
                                         uint32_t n_refs, const gis_params *params, gis_ref_randstrobe *out,
                                         uint64_t capacity, uint64_t *n_out) {
    API_BEGIN(ctx)
    if (!params || !n_out) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null argument");
    if (int rc = check_contigs(ctx, seqs, lens, n_refs)) return rc;
    *n_out = 0;
    uint64_t total = 0;
    for (uint32_t j = 0; j < n_refs; j++) total += lens[j];
    if (total == 0) return GIS_OK;
    uint4 *recs = nullptr;
    uint64_t n = 0;
    try {
        device_seed_reference(seqs, lens, n_refs, *params, &recs, &n, ctx->stream);
    } catch (const BuildError &e) {
        throw HipError{e.e, e.what};
    }
    DevBuf keep{recs, n * sizeof(gis_ref_randstrobe)};
    *n_out = n;
    if (n > capacity || (n && !out)) {
        release(keep);
        return fail(ctx, GIS_ERR_CAPACITY, "record capacity " + std::to_string(capacity) + " < " + std::to_string(n));
    }
    const hipError_t e = n ? hipMemcpy(out, recs, n * sizeof(gis_ref_randstrobe), hipMemcpyDeviceToHost) : hipSuccess;
    release(keep);
    HIP_CHECK(e);
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

// .sti v2 (step2/index.cu:73-134) mapped into memory and copied from the
// mapping straight into HBM
extern "C" int gis_index_read_sti(gis_ctx *ctx, const char *path, const gis_params *expected) {
    API_BEGIN(ctx)
    if (!path) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null path");
    MappedSti m;
    std::string err;
    const int rc = map_sti(path, expected, m, err);
    if (rc) return fail(ctx, rc, err);
    try {
        DevIndexStore &ix = ctx->ix;
        drop_table(ix);
        ensure<gis_ref_randstrobe>(ix.rs, m.n);
        ensure<uint64_t>(ix.starts, (1ull << m.bits) + 1);
        const uint64_t chunk = 256ull << 20;
        const char *src = reinterpret_cast<const char *>(m.rs);
        char *dst = static_cast<char *>(ix.rs.p);
        for (uint64_t o = 0, tot = m.n * sizeof(gis_ref_randstrobe); o < tot; o += chunk)
            HIP_CHECK(hipMemcpy(dst + o, src + o, std::min(chunk, tot - o), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(ix.starts.p, m.starts, ((1ull << m.bits) + 1) * 8, hipMemcpyHostToDevice));
        // the file's table must be one the lookups can trust
        const char *why = nullptr;
        try {
            why = device_validate_index(static_cast<const uint4 *>(ix.rs.p), m.n,
                                        static_cast<const uint64_t *>(ix.starts.p), m.bits, ctx->stream);
        } catch (const BuildError &e) {
            throw HipError{e.e, e.what};
        }
        if (why) {
            drop_table(ix);
            unmap_sti(m);
            return fail(ctx, GIS_ERR_INVALID_INDEX, std::string("Index file has incorrect format: ") + why);
        }
        ix.n = m.n;
        ix.bits = m.bits;
        ix.filter_cutoff = m.filter_cutoff;
        ix.params = m.params;
        ix.distinct = ix.once = ix.high = ix.mid = ix.index_cutoff = 0;
        build_lines(ctx);
        ix.loaded = true;
    } catch (...) {
        unmap_sti(m);
        throw;
    }
    unmap_sti(m);
    return GIS_OK;
    API_END(ctx)
}

// .sti v2 written from HBM in chunks (no whole host copy of the table)
extern "C" int gis_index_write_sti(gis_ctx *ctx, const char *path) {
    API_BEGIN(ctx)
    if (!path) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null path");
    if (!ctx->ix.loaded) return fail(ctx, GIS_ERR_STATE, "no index loaded");
    const DevIndexStore &ix = ctx->ix;
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return fail(ctx, GIS_ERR_IO, std::string(path) + ": " + std::strerror(errno));
    bool ok = true;
    auto put = [&](const void *d, size_t n) { ok = ok && std::fwrite(d, 1, n, fp) == n; };
    put("STI\1", 4);
    const uint32_t version = 2;
    put(&version, 4);
    // a reserved chunk of 8 zero bytes, as the reference writes it
    const uint64_t reserved_size = 8, reserved = 0;
    put(&reserved_size, 8);
    put(&reserved, 8);
    const gis_params &p = ix.params;
    const int32_t v32[9] = {(int32_t)ix.filter_cutoff, ix.bits, p.canonical_read_length, p.k, p.s, p.l, p.u,
                            (int32_t)p.q, p.max_dist};
    put(v32, sizeof v32);
    put(&ix.n, 8);
    std::vector<char> buf;
    const uint64_t chunk = 256ull << 20;
    const uint64_t tot = ix.n * sizeof(gis_ref_randstrobe);
    try {
        buf.resize((size_t)std::min(chunk, std::max<uint64_t>(tot, 8)));
        for (uint64_t o = 0; o < tot && ok; o += chunk) {
            const uint64_t k = std::min(chunk, tot - o);
            HIP_CHECK(hipMemcpy(buf.data(), static_cast<const char *>(ix.rs.p) + o, k, hipMemcpyDeviceToHost));
            put(buf.data(), k);
        }
        const uint64_t nb = (1ull << ix.bits) + 1;
        put(&nb, 8);
        std::vector<uint64_t> st(nb);
        HIP_CHECK(hipMemcpy(st.data(), ix.starts.p, nb * 8, hipMemcpyDeviceToHost));
        put(st.data(), nb * 8);
    } catch (...) {
        std::fclose(fp);
        throw;
    }
    if (std::fclose(fp) != 0) ok = false;
    if (!ok) return fail(ctx, GIS_ERR_IO, std::string(path) + ": write failed");
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
    out->device_bytes = ix.rs.bytes + ix.starts.bytes + ix.lines.bytes;
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

extern "C" int gis_fused_config_get(gis_ctx *ctx, gis_fused_config *out) {
    if (!ctx || !out) return GIS_ERR_BAD_PARAMETER;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    *out = ctx->fcfg;
    return GIS_OK;
}

extern "C" int gis_fused_config_set(gis_ctx *ctx, const gis_fused_config *cfg) {
    if (!ctx || !cfg) return GIS_ERR_BAD_PARAMETER;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    const gis_fused_config c = *cfg;
    if (c.waves_per_block != 2 && c.waves_per_block != 4 && c.waves_per_block != 8)
        return fail(ctx, GIS_ERR_BAD_PARAMETER, "waves_per_block must be 2, 4 or 8");
    if (c.length_split != 0 && c.length_split != 1)
        return fail(ctx, GIS_ERR_BAD_PARAMETER, "length_split must be 0 or 1");
    if (c.queue != 0 && c.queue != 1) return fail(ctx, GIS_ERR_BAD_PARAMETER, "queue must be 0 or 1");
    if (c.queue == 1 && (c.chunk < 1 || c.chunk > 4096))
        return fail(ctx, GIS_ERR_BAD_PARAMETER, "chunk must be in 1..4096");
    if (c.blocks_per_cu < 0 || c.blocks_per_cu > 64)
        return fail(ctx, GIS_ERR_BAD_PARAMETER, "blocks_per_cu must be in 0..64");
    if (c.queue == 1 && c.blocks_per_cu == 0)
        return fail(ctx, GIS_ERR_BAD_PARAMETER, "a dynamic queue needs a persistent grid (blocks_per_cu > 0)");
    ctx->fcfg = c;
    if (c.queue == 0) ctx->fcfg.chunk = 0;
    return GIS_OK;
}
