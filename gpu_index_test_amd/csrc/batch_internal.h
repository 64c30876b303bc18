// The device batch shared by the C-ABI batch entry points (batch.cpp) and
// the serving loop (server.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "driver_internal.h"

namespace gis {

// Grow-only pinned host array: copies between it and HBM are DMA transfers
// that run asynchronously to the host.
template <typename T>
struct Pinned {
    T *p = nullptr;
    size_t cap = 0, n = 0;
    Pinned() = default;
    Pinned(const Pinned &) = delete;
    Pinned &operator=(const Pinned &) = delete;
    Pinned(Pinned &&o) noexcept : p(o.p), cap(o.cap), n(o.n) { o.p = nullptr, o.cap = o.n = 0; }
    Pinned &operator=(Pinned &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p, cap = o.cap, n = o.n;
            o.p = nullptr, o.cap = o.n = 0;
        }
        return *this;
    }
    ~Pinned() { release(); }
    void reserve(size_t want) {
        if (want <= cap) return;
        const size_t c = std::max(want, cap + cap / 2);
        T *q = nullptr;
        if (hipHostMalloc((void **)&q, std::max<size_t>(c, 1) * sizeof(T)) != hipSuccess)
            throw HipError{hipErrorOutOfMemory, "hipHostMalloc of " + std::to_string(c * sizeof(T)) + " bytes"};
        if (p && n) std::memcpy(q, p, n * sizeof(T));
        if (p) (void)hipHostFree(p);
        p = q;
        cap = c;
    }
    void resize(size_t m) {
        reserve(m);
        n = m;
    }
    void assign(size_t m, T v) {
        resize(m);
        std::fill(p, p + m, v);
    }
    void assign(const T *a, size_t m) {
        resize(m);
        if (m) std::memcpy(p, a, m * sizeof(T));
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = n = 0;
    }
    T *data() { return p; }
    const T *data() const { return p; }
    size_t size() const { return n; }
    T &operator[](size_t i) { return p[i]; }
    const T &operator[](size_t i) const { return p[i]; }
};

// Device intermediates and results of a run. A workspace may be shared by
// several batches that run one at a time (each run's results must be read
// before another batch runs on it); each serving slot has its own.
struct Workspace {
    DevBuf long_scratch;
    DevBuf rs_cnt, rs_nfwd, rs_hash, rs_span, sy_cnt, sy_hash, sy_pos;
    DevBuf pos, hcnt, rcnt, lk_bytes, lk_stream;
    DevBuf hit_total, rhit_total, frac, flags, nam_cnt, hit_off, rhit_off;
    DevBuf hits, rhits, keys, rkeys, nam_ws, rnam_ws, nam_fin, rnam_fin;
    DevBuf nam_slots, nam_off, nams_out, scan_tmp, ctl, spill_ids, gwork, split_list, sel_ids;
    uint64_t hit_cap = 0, rhit_cap = 0, out_cap = 0;
    bool merge_scratch = false;  // keys/nam_ws/gwork allocated (reads with many hits need them)
    // results of the last run
    uint64_t tot_nams = 0;
    uint32_t n_spill = 0;
    const gis_batch *owner = nullptr;  // the batch whose run the results belong to
    // the batch whose run was launched (gis_batch_launch) and not yet waited
    // for: the workspace takes no other run meanwhile
    const gis_batch *in_flight = nullptr;
    // a stream created for this workspace (destroyed with it), else null
    hipStream_t stream = nullptr;
    // pinned landing area for the end-of-run readback, and the device's
    // address of it (a kernel writes it)
    DevCtl *h_ctl = nullptr;
    DevCtl *h_ctl_dev = nullptr;
    int refs = 0;
    PairWork *pairs = nullptr;    // the pair step on this workspace's NAMs (pairs.cpp)
};

}  // namespace gis

struct gis_batch {
    using DevBuf = gis::DevBuf;
    using DevBatch = gis::DevBatch;
    gis_ctx *ctx = nullptr;
    // resident reads and what the host derived from their lengths
    uint32_t n_reads = 0;
    uint64_t span = 0;  // bases
    bool want_sy = false;
    bool any_long = false;
    // pinned, so that their copies to HBM run while the host works on
    gis::Pinned<uint64_t> h_rs_off, h_sy_off, h_long_off, h_read_off;
    DevBuf seqs, read_off, read_len, rs_off, sy_off, long_off, queue;
    gis::Pinned<uint32_t> h_queue;  // read ids by length, longest first
    uint32_t n_long = 0;            // queue entries longer than the short bin (160 bp)
    bool queue_ok = false;          // the queue was staged with the reads (fused runs need it)
    uint64_t tot_rs_cap = 0;
    // seeding parameters of a run that needs no index (else the index's)
    const gis_params *seed_params = nullptr;
    gis::Workspace *ws = nullptr;
    int last_mode = GIS_MODE_PLAIN;
    bool last_graph = false;
    bool ran = false;
    hipEvent_t ev[16] = {};
    int n_ev = 0;
    std::vector<std::string> ev_names;
    // captured launch sequences, one per configuration (buffers, mode, cutoffs)
    struct GraphEntry {
        DevBatch key;
        int cfg[11];
        uint64_t epoch;  // the index table the graph was captured against
        hipGraph_t graph;
        hipGraphExec_t exec;
        uint64_t used;
    };
    std::vector<GraphEntry> graphs;
    uint64_t graph_clock = 0;
    // the stream the batch runs on: the context's, or its workspace's own
    // (a serving slot, gis_batch_create_async and the batches sharing it)
    hipStream_t stream = nullptr;
    // the options of the launched run gis_batch_wait completes
    gis_map_options pending_opt{};
};

namespace gis {

enum { STOP_SEED = 0, STOP_HITS = 1, STOP_NONE = 2 };

// the launch configuration of one run
struct RunOpts {
    int mode;          // GIS_MODE_*
    int rescue_mode;   // 0 the caller's decision, 1 never, 2 always
    uint32_t rcut, hard;
    int stop;          // STOP_*
    bool graph;        // replay a captured graph
};

gis_batch *batch_new(gis_ctx *ctx, Workspace *share);  // share = nullptr: a workspace of its own
gis_batch *batch_new_stream(gis_ctx *ctx);  // a batch with a stream and a workspace of its own
// device bytes the batch's inputs and workspace hold
uint64_t batch_device_bytes(gis_batch *b);
// free the batch's device arrays and captured graphs (after its stream
// drains); the next run allocates them again. The workspace must be the
// batch's own.
void batch_trim(gis_batch *b);
// One host-to-device or device-to-host copy.
struct CopyItem {
    void *dst;
    const void *src;
    size_t bytes;
};

// Stage reads into the batch without waiting: the host derives the
// per-read arrays into the batch's pinned arrays, then the copies to HBM are
// enqueued on the batch's stream (after `before`, when given, is recorded
// there), or, with `defer`, appended to it for the caller to issue before the
// run. Device offsets are offsets[r] - base and the bases copied are
// seqs[base .. max(offsets[r] + lens[r])); seqs and lens are read by the copies
// as they run, so they stay valid until then. Without want_queue the fused
// form's work queue is not built (plain runs only).
int batch_stage(gis_batch *b, const char *seqs, const uint64_t *offsets, const uint32_t *lens,
                uint32_t n, bool want_sy, uint64_t base, hipEvent_t before,
                std::vector<CopyItem> *defer = nullptr, bool want_queue = true);
// batch_stage with base 0, then wait for the copies
int batch_upload(gis_batch *b, const char *seqs, const uint64_t *offsets, const uint32_t *lens,
                 uint32_t n, bool want_sy);
RunOpts run_opts(const gis_batch *b, const gis_map_options &o, int rescue_mode, int stop);
// enqueue one run on the batch's stream without waiting
void batch_launch(gis_batch *b, const RunOpts &r);
// the launched run has finished on the device
bool batch_ready(const gis_batch *b);
// wait for the launched run, grow what overflowed and run again until the
// results are complete
void batch_complete(gis_batch *b, const RunOpts &r);
void batch_run(gis_batch *b, const gis_map_options &o, int rescue_mode, int stop);
int batch_download(gis_batch *b, gis_nam_out *out);
gis_map_options opts_or_default(const gis_map_options *opt);

}  // namespace gis
