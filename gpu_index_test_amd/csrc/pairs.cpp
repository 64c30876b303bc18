// C ABI: the NAM consumer of paired reads (kernels in pair_kernels.hip).
//
// A pair step is one sequence of launches on the batch's stream: reference
// order, pair types, per-type lists and, with GIS_PAIRS_JOINT, the joint NAM
// pairs. The joint entries are placed by a scan of per-pair upper bounds, so
// the host reads that total once to size the entry array; a second
// synchronisation at the end reads the type counts, the entry total and the
// error bits.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "batch_internal.h"
#include "pairs.h"

using namespace gis;

namespace gis {

struct PairWork {
    DevBuf in_off, in_nams;  // lists uploaded by gis_pairs_from_nams
    DevBuf ord, types, drop, mapq, tcnt, tpos, tids, ttot, jbound, jboff, jtmp, jcnt, joff, jout, err, scan_tmp;
    uint64_t *h = nullptr;  // pinned: [0..5] type starts, [6] joint bound total, [7] joint total, [8] error bits
    const void *owner = nullptr;  // the batch (or context) whose lists the results belong to
    const uint64_t *dev_off = nullptr;
    uint32_t n_pairs = 0;
    uint64_t n_nams = 0, n_joint = 0;
    uint32_t type_counts[5] = {};
    bool joint = false;
};

void pairwork_free(PairWork *w) {
    if (!w) return;
    DevBuf *all[] = {&w->in_off, &w->in_nams, &w->ord,  &w->types, &w->drop, &w->mapq,
                     &w->tcnt,   &w->tpos,    &w->tids, &w->ttot,  &w->jbound, &w->jboff,
                     &w->jtmp,   &w->jcnt,    &w->joff, &w->jout,  &w->err,  &w->scan_tmp};
    for (DevBuf *x : all) release(*x);
    if (w->h) (void)hipHostFree(w->h);
    delete w;
}

static PairWork *pairwork_get(PairWork *&slot) {
    if (!slot) {
        slot = new PairWork();
        if (hipHostMalloc((void **)&slot->h, 16 * sizeof(uint64_t), hipHostMallocDefault) != hipSuccess) {
            slot->h = nullptr;
            throw HipError{hipErrorOutOfMemory, "hipHostMalloc of the pair step's readback area"};
        }
    }
    return slot;
}

static int check_params(gis_ctx *ctx, const gis_pair_params &pp) {
    if (pp.max_tries < 1 || pp.max_tries > PAIR_MAX_TRIES)
        return fail(ctx, GIS_ERR_BAD_PARAMETER,
                    "max_tries must be in 1 .. " + std::to_string(PAIR_MAX_TRIES));  // step2/aln.hpp:71-73
    if (!std::isfinite(pp.dropoff_threshold) || !std::isfinite(pp.mu) || !std::isfinite(pp.sigma))
        return fail(ctx, GIS_ERR_BAD_PARAMETER, "pair parameters must be finite");
    return GIS_OK;
}

// the pair step on device lists off[2n+1] / in
static int pairs_run(gis_ctx *ctx, PairWork &w, hipStream_t st, const uint64_t *off, const gis_nam *in,
                     uint64_t n_nams, uint32_t n_pairs, int grouped, const gis_pair_params &pp) {
    const size_t nb = pair_blocks(n_pairs);
    PairDev P{};
    P.off = off;
    P.in = in;
    P.ord = ensure<gis_nam>(w.ord, n_nams);
    P.n_pairs = n_pairs;
    P.grouped = grouped;
    P.types = ensure<uint8_t>(w.types, n_pairs);
    P.drop = ensure<float>(w.drop, 2ull * n_pairs);
    P.mapq = ensure<uint8_t>(w.mapq, 2ull * n_pairs);
    P.tcnt = ensure<uint32_t>(w.tcnt, 5 * nb);
    P.tpos = ensure<uint64_t>(w.tpos, 5 * nb + 1);
    P.tids = ensure<uint32_t>(w.tids, n_pairs);
    P.ttot = ensure<uint64_t>(w.ttot, 6);
    P.nblk = (uint32_t)nb;
    P.err = ensure<unsigned>(w.err, 1);
    P.dropoff = pp.dropoff_threshold;
    P.mu = pp.mu;
    P.sigma = pp.sigma;
    P.max_tries = pp.max_tries;
    P.joint = (pp.flags & GIS_PAIRS_JOINT) ? 1 : 0;
    if (P.joint) {
        P.jbound = ensure<uint32_t>(w.jbound, n_pairs);
        P.jboff = ensure<uint64_t>(w.jboff, n_pairs + 1ull);
        P.jcnt = ensure<uint32_t>(w.jcnt, n_pairs);
        P.joff = ensure<uint64_t>(w.joff, n_pairs + 1ull);
    }
    uint64_t *tmp = ensure<uint64_t>(w.scan_tmp, scan_tmp_words(std::max<uint64_t>(5 * nb, n_pairs)));
    w.owner = nullptr;
    HIP_CHECK(hipMemsetAsync(P.err, 0, sizeof(unsigned), st));
    launch_pairs_classify(P, tmp, st);
    uint64_t *h = w.h;
    if (P.joint) {
        HIP_CHECK(hipMemcpyAsync(&h[6], P.jboff + n_pairs, 8, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        const uint64_t bound = h[6];
        P.jtmp = ensure<gis_nam_pair>(w.jtmp, bound);
        P.jout = ensure<gis_nam_pair>(w.jout, bound);
        launch_pairs_joint(P, tmp, st);
        HIP_CHECK(hipMemcpyAsync(&h[7], P.joff + n_pairs, 8, hipMemcpyDeviceToHost, st));
    }
    HIP_CHECK(hipMemcpyAsync(h, P.ttot, 6 * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    unsigned *herr = reinterpret_cast<unsigned *>(&h[8]);
    HIP_CHECK(hipMemcpyAsync(herr, P.err, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    if (*herr & PAIR_ERR_DUPLICATE)
        return fail(ctx, GIS_ERR_BAD_PARAMETER,
                    "a NAM list holds two NAMs with the same (is_rc, ref_id, query_start, ref_start)");
    if (*herr & PAIR_ERR_JOINT) return fail(ctx, GIS_ERR_STATE, "joint NAM pairs exceeded their bound");
    for (int t = 0; t < 5; t++) w.type_counts[t] = (uint32_t)(h[t + 1] - h[t]);
    w.n_pairs = n_pairs;
    w.n_nams = n_nams;
    w.n_joint = P.joint ? h[7] : 0;
    w.joint = P.joint != 0;
    w.dev_off = off;
    return GIS_OK;
}

static int pairs_download(gis_ctx *ctx, PairWork &w, hipStream_t st, gis_pair_out *out) {
    if (!out) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null output");
    const uint32_t n = w.n_pairs;
    out->n_nams = w.n_nams;
    out->n_joint = w.n_joint;
    std::memcpy(out->type_counts, w.type_counts, sizeof w.type_counts);
    if (out->nams && out->nam_capacity < w.n_nams)
        return fail(ctx, GIS_ERR_CAPACITY,
                    "NAM capacity " + std::to_string(out->nam_capacity) + " < " + std::to_string(w.n_nams));
    if (out->joint && out->joint_capacity < w.n_joint)
        return fail(ctx, GIS_ERR_CAPACITY, "joint capacity " + std::to_string(out->joint_capacity) + " < " +
                                               std::to_string(w.n_joint));
    if (out->nam_offsets)
        HIP_CHECK(hipMemcpyAsync(out->nam_offsets, w.dev_off, 8ull * (2ull * n + 1), hipMemcpyDeviceToHost, st));
    if (out->nams && w.n_nams)
        HIP_CHECK(hipMemcpyAsync(out->nams, w.ord.p, sizeof(gis_nam) * w.n_nams, hipMemcpyDeviceToHost, st));
    if (out->types && n) HIP_CHECK(hipMemcpyAsync(out->types, w.types.p, n, hipMemcpyDeviceToHost, st));
    if (out->type_ids && n)
        HIP_CHECK(hipMemcpyAsync(out->type_ids, w.tids.p, 4ull * n, hipMemcpyDeviceToHost, st));
    if (out->top_dropoff && n)
        HIP_CHECK(hipMemcpyAsync(out->top_dropoff, w.drop.p, 8ull * n, hipMemcpyDeviceToHost, st));
    if (out->mapq && n) HIP_CHECK(hipMemcpyAsync(out->mapq, w.mapq.p, 2ull * n, hipMemcpyDeviceToHost, st));
    if (out->joint_offsets) {
        if (w.joint)
            HIP_CHECK(hipMemcpyAsync(out->joint_offsets, w.joff.p, 8ull * (n + 1ull), hipMemcpyDeviceToHost, st));
        else std::memset(out->joint_offsets, 0, 8ull * (n + 1ull));
    }
    if (out->joint && w.n_joint)
        HIP_CHECK(hipMemcpyAsync(out->joint, w.jout.p, sizeof(gis_nam_pair) * w.n_joint, hipMemcpyDeviceToHost,
                                 st));
    HIP_CHECK(hipStreamSynchronize(st));
    return GIS_OK;
}

}  // namespace gis

extern "C" void gis_pair_params_default(gis_pair_params *p) {
    if (!p) return;
    p->dropoff_threshold = 0.5f;
    p->max_tries = 20;
    p->mu = 300.0f;
    p->sigma = 100.0f;
    p->flags = 0;
}

extern "C" int gis_batch_pairs(gis_batch *b, const gis_pair_params *pp) {
    if (!b) return GIS_ERR_BAD_PARAMETER;
    gis_ctx *ctx = b->ctx;
    API_BEGIN(ctx)
    gis_pair_params p;
    if (pp) p = *pp;
    else gis_pair_params_default(&p);
    if (int rc = check_params(ctx, p)) return rc;
    if (!b->ran) return fail(ctx, GIS_ERR_STATE, "batch has not been run");
    if (b->ws->owner != b)
        return fail(ctx, GIS_ERR_STATE, "another batch sharing the workspace ran since this batch's run");
    if (b->n_reads % 2)
        return fail(ctx, GIS_ERR_BAD_PARAMETER,
                    "a paired batch holds 2 n reads (mate 1 of pair p is read p, mate 2 is read n + p)");
    PairWork *w = pairwork_get(b->ws->pairs);
    int rc = pairs_run(ctx, *w, b->stream, static_cast<const uint64_t *>(b->ws->nam_off.p),
                       static_cast<const gis_nam *>(b->ws->nams_out.p), b->ws->tot_nams, b->n_reads / 2, 1, p);
    if (rc == GIS_OK) w->owner = b;
    return rc;
    API_END(ctx)
}

extern "C" int gis_batch_pairs_download(gis_batch *b, gis_pair_out *out) {
    if (!b) return GIS_ERR_BAD_PARAMETER;
    gis_ctx *ctx = b->ctx;
    API_BEGIN(ctx)
    PairWork *w = b->ws->pairs;
    if (!w || w->owner != b || b->ws->owner != b)
        return fail(ctx, GIS_ERR_STATE, "no pair step of this batch's last run");
    return pairs_download(ctx, *w, b->stream, out);
    API_END(ctx)
}

extern "C" int gis_pairs_from_nams(gis_ctx *ctx, const uint64_t *nam_offsets, const gis_nam *nams,
                                   uint32_t n_pairs, const gis_pair_params *pp, gis_pair_out *out) {
    API_BEGIN(ctx)
    gis_pair_params p;
    if (pp) p = *pp;
    else gis_pair_params_default(&p);
    if (int rc = check_params(ctx, p)) return rc;
    if (!nam_offsets) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null NAM offsets");
    const uint64_t n_lists = 2ull * n_pairs;
    if (nam_offsets[0] != 0) return fail(ctx, GIS_ERR_BAD_PARAMETER, "NAM offsets must start at 0");
    for (uint64_t r = 0; r < n_lists; r++)
        if (nam_offsets[r + 1] < nam_offsets[r])
            return fail(ctx, GIS_ERR_BAD_PARAMETER, "NAM offsets must not decrease");
    const uint64_t total = nam_offsets[n_lists];
    if (total && !nams) return fail(ctx, GIS_ERR_BAD_PARAMETER, "null NAMs");
    PairWork *w = pairwork_get(ctx->pairs);
    uint64_t *doff = ensure<uint64_t>(w->in_off, n_lists + 1);
    gis_nam *dn = ensure<gis_nam>(w->in_nams, total);
    HIP_CHECK(hipMemcpyAsync(doff, nam_offsets, 8 * (n_lists + 1), hipMemcpyHostToDevice, ctx->stream));
    if (total) HIP_CHECK(hipMemcpyAsync(dn, nams, sizeof(gis_nam) * total, hipMemcpyHostToDevice, ctx->stream));
    int rc = pairs_run(ctx, *w, ctx->stream, doff, dn, total, n_pairs, 0, p);
    if (rc) return rc;
    w->owner = ctx;
    return pairs_download(ctx, *w, ctx->stream, out);
    API_END(ctx)
}
