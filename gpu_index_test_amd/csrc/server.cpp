// C ABI: the serving loop. Requests (sets of reads) are queued by the
// caller; slot threads admit them into device slots by an HBM budget,
// coalesces the queued requests into one launch sequence per slot (a step,
// replayed from the slot's captured hipGraph for its padded size), keeps up
// to `slots` steps in flight on separate streams, and hands each request its
// per-read NAMs when its step completes. Requests that arrive while steps
// are on the GPU join the next step on a free slot (continuous batching);
// a request's NAMs do not depend on the requests it was batched with. A
// request larger than one step (max_step_reads, or the HBM budget) is cut
// into parts that run as steps of their own and are put back together when
// the caller waits for it. Admission charges each slot the device memory its
// batch really holds: a running step the larger of what its batch held and
// the step's estimate, an idle slot what its batch still holds; idle slots
// give theirs back when a step would not fit otherwise.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch_internal.h"

using namespace gis;
using Clock = std::chrono::steady_clock;

namespace {

struct Request {
    uint64_t id = 0;
    std::vector<char> seqs;
    std::vector<uint64_t> offs;
    std::vector<uint32_t> lens;
    uint64_t bytes = 0;  // estimated device bytes while in flight
    Clock::time_point t_submit, t_done;
    bool done = false;
    int status = GIS_OK;
    std::string err;
    std::vector<uint64_t> nam_off;
    std::vector<gis_nam> nams;
    std::vector<float> frac;
    std::vector<uint8_t> flags;
    // a request cut into steps: its parts in read order (each runs as a
    // request of its own) and the parts still running; a part's parent
    Request *parent = nullptr;
    std::vector<Request *> parts;
    size_t parts_left = 0;
    bool claimed = false;  // a caller is waiting for it
    Request() = default;
    Request(const Request &) = delete;
    Request &operator=(const Request &) = delete;
    ~Request() {
        for (Request *p : parts) delete p;
    }
};

struct Slot {
    gis_batch *b = nullptr;
    std::vector<Request *> reqs;
    std::vector<uint32_t> first;  // first read of each request in the step
    uint32_t n_reads = 0;         // reads of the step before padding
    bool busy = false;            // a step is running
    uint64_t held = 0;            // device bytes the slot's batch holds between steps
    uint64_t charge = 0;          // bytes charged while its step runs
    RunOpts r{};
    std::thread th;
    // pinned host staging of the slot's thread
    Pinned<char> st_seqs;
    Pinned<uint64_t> st_offs;
    Pinned<uint32_t> st_lens;
    Pinned<uint64_t> h_off;
    Pinned<gis_nam> h_nams;
    Pinned<float> h_frac;
    Pinned<uint8_t> h_flags;
    void release() {
        st_seqs.release();
        st_offs.release();
        st_lens.release();
        h_off.release();
        h_nams.release();
        h_frac.release();
        h_flags.release();
    }
};

// device bytes a request holds while in flight: randstrobe slots and
// lookup results (~64 B per base); per read its 128-hit region of the hit
// array (2 KB), the merge's NAM staging beside it (5.6 KB), merge keys and
// per-read arrays (~10 KB per read)
uint64_t estimate_bytes(uint64_t reads, uint64_t bases) { return 64 * bases + 10240 * reads + 4096; }

// padded step sizes, so that a slot replays one of a few captured graphs
uint32_t padded(uint32_t n) {
    uint32_t p = 256;
    while (p < n) p <<= 1;
    return p;
}

}  // namespace

struct gis_server {
    gis_ctx *ctx = nullptr;
    gis_server_config cfg{};
    std::vector<Slot> slots;  // sized once; each slot runs on a thread of its own
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> pending;
    std::unordered_map<uint64_t, Request *> reqs;
    uint64_t next_id = 1;
    int n_waiters = 0;
    bool stop = false;
    gis_server_stats stats{};
    Clock::time_point busy_since;
    int n_busy = 0;
};

namespace {

// (the slot's thread marks the requests done under the server lock afterwards)
void fail_step(gis_server *s, Slot &sl, const std::string &msg) {
    (void)s;
    for (Request *q : sl.reqs) {
        q->status = GIS_ERR_HIP;
        q->err = msg;
        q->t_done = Clock::now();
    }
}

// stage, upload and launch the requests of a slot (on the slot's thread)
void launch_step(gis_server *s, Slot &sl) {
    const uint32_t n = sl.n_reads, np = padded(n);
    uint64_t bases = 0;
    for (Request *q : sl.reqs) bases += q->seqs.size();
    sl.st_seqs.resize(bases + 1);
    sl.st_offs.resize(np);
    sl.st_lens.resize(np);
    sl.first.clear();
    uint32_t r = 0;
    uint64_t at = 0;
    for (Request *q : sl.reqs) {
        sl.first.push_back(r);
        if (!q->seqs.empty()) std::memcpy(sl.st_seqs.data() + at, q->seqs.data(), q->seqs.size());
        for (size_t i = 0; i < q->lens.size(); i++, r++) {
            sl.st_offs[r] = at + q->offs[i];
            sl.st_lens[r] = q->lens[i];
        }
        at += q->seqs.size();
    }
    for (; r < np; r++) {  // empty reads: no randstrobes, no NAMs
        sl.st_offs[r] = at;
        sl.st_lens[r] = 0;
    }
    const int rc = batch_upload(sl.b, sl.st_seqs.data(), sl.st_offs.data(), sl.st_lens.data(), np, false);
    if (rc) throw HipError{hipErrorInvalidValue, "server upload failed"};
    sl.r = run_opts(sl.b, s->cfg.options, 0, STOP_NONE);
    batch_launch(sl.b, sl.r);
}

// wait for the slot's step and split its results into its requests (on the
// slot's thread)
void complete_step(gis_server *s, Slot &sl) {
    (void)s;
    batch_complete(sl.b, sl.r);
    const uint32_t np = sl.b->n_reads;
    sl.h_off.resize(np + 1);
    sl.h_nams.resize(std::max<uint64_t>(1, sl.b->ws->tot_nams));
    sl.h_frac.resize(np);
    sl.h_flags.resize(np);
    gis_nam_out o{sl.h_off.data(), sl.h_nams.data(), sl.h_nams.size(), sl.h_frac.data(), sl.h_flags.data(), 0};
    if (batch_download(sl.b, &o)) throw HipError{hipErrorInvalidValue, "server download failed"};
    const Clock::time_point now = Clock::now();
    for (size_t j = 0; j < sl.reqs.size(); j++) {
        Request *q = sl.reqs[j];
        const uint32_t f = sl.first[j], k = (uint32_t)q->lens.size();
        const uint64_t a = sl.h_off[f], e = sl.h_off[f + k];
        q->nam_off.resize(k + 1);
        for (uint32_t i = 0; i <= k; i++) q->nam_off[i] = sl.h_off[f + i] - a;
        q->nams.assign(sl.h_nams.data() + a, sl.h_nams.data() + e);
        q->frac.assign(sl.h_frac.data() + f, sl.h_frac.data() + f + k);
        q->flags.assign(sl.h_flags.data() + f, sl.h_flags.data() + f + k);
        q->t_done = now;
    }
}

// Device bytes charged to all slots but `self`, split into running steps
// and idle slots.
void charged(const gis_server *s, size_t self, uint64_t &running, uint64_t &idle) {
    running = idle = 0;
    for (size_t j = 0; j < s->slots.size(); j++) {
        if (j == self) continue;
        if (s->slots[j].busy) running += s->slots[j].charge;
        else idle += s->slots[j].held;
    }
}

// Can slot si run a step of estimated bytes e now? Idle slots' arrays are
// given back (largest first) when only they stand in the way. Called with
// the server lock held.
bool admit(gis_server *s, size_t si, uint64_t e) {
    uint64_t running, idle;
    charged(s, si, running, idle);
    const uint64_t mine = std::max(s->slots[si].held, e);
    if (running + idle + mine <= s->cfg.hbm_budget) return true;
    if (running + mine > s->cfg.hbm_budget) return false;
    while (running + idle + mine > s->cfg.hbm_budget) {
        Slot *big = nullptr;
        for (size_t j = 0; j < s->slots.size(); j++) {
            Slot &o = s->slots[j];
            if (j != si && !o.busy && o.held && (!big || o.held > big->held)) big = &o;
        }
        if (!big) break;
        try {
            batch_trim(big->b);
        } catch (const HipError &) {
            return false;
        }
        idle -= big->held;
        big->held = batch_device_bytes(big->b);
        idle += big->held;
    }
    return running + idle + mine <= s->cfg.hbm_budget;
}

// One thread per slot: take the queued requests that fit (arrival order,
// max_step_reads, HBM budget), stage and launch them as one step, wait for
// it on the slot's stream, and complete the requests. Slots run their host
// work and their device work independently, so steps overlap on the GPU and
// with each other's copies.
void slot_main(gis_server *s, size_t si) {
    (void)hipSetDevice(s->ctx->device);
    Slot &sl = s->slots[si];
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_work.wait(lk, [&] {
            return s->stop || (!s->pending.empty() && admit(s, si, s->pending.front()->bytes));
        });
        if (s->stop) break;
        uint32_t n = 0;
        uint64_t bytes = 0;
        std::vector<Request *> take;
        while (!s->pending.empty()) {
            Request *q = s->pending.front();
            const uint32_t k = (uint32_t)q->lens.size();
            if (!take.empty() && (n + k > s->cfg.max_step_reads || !admit(s, si, bytes + q->bytes))) break;
            take.push_back(q);
            n += k;
            bytes += q->bytes;
            s->pending.pop_front();
        }
        sl.reqs = std::move(take);
        sl.n_reads = n;
        sl.busy = true;
        sl.charge = std::max(sl.held, bytes);
        {
            uint64_t running, idle;
            charged(s, s->slots.size(), running, idle);
            s->stats.peak_bytes = std::max(s->stats.peak_bytes, running + idle);
        }
        if (s->n_busy++ == 0) s->busy_since = Clock::now();
        s->stats.steps++;
        lk.unlock();
        uint64_t held = 0;
        try {
            launch_step(s, sl);
            complete_step(s, sl);
        } catch (const HipError &e) {
            fail_step(s, sl, e.what + ": " + hipGetErrorString(e.e));
            (void)hipStreamSynchronize(sl.b->stream);
        } catch (const std::bad_alloc &) {
            fail_step(s, sl, "host allocation failed");
            (void)hipStreamSynchronize(sl.b->stream);
        }
        // a slot keeps its arrays for the next step up to its share of the
        // budget
        try {
            held = batch_device_bytes(sl.b);
            if (held > s->cfg.hbm_budget / s->slots.size()) {
                batch_trim(sl.b);
                held = batch_device_bytes(sl.b);
            }
        } catch (const HipError &) {
        }
        lk.lock();
        for (Request *q : sl.reqs) {
            q->done = true;
            if (q->parent && --q->parent->parts_left == 0) q->parent->done = true;
        }
        sl.reqs.clear();
        sl.busy = false;
        sl.held = held;
        sl.charge = 0;
        if (--s->n_busy == 0)
            s->stats.busy_s += std::chrono::duration<double>(Clock::now() - s->busy_since).count();
        s->cv_done.notify_all();
        s->cv_work.notify_all();  // budget freed
    }
}

Request *make_request(gis_server *s, const char *seqs, const uint64_t *offsets, const uint32_t *lens,
                      uint32_t n_reads) {
    Request *q = new Request();
    uint64_t bases = 0;
    q->offs.resize(n_reads);
    q->lens.assign(lens, lens + n_reads);
    for (uint32_t i = 0; i < n_reads; i++) bases += lens[i];
    q->seqs.reserve(bases);
    for (uint32_t i = 0; i < n_reads; i++) {
        q->offs[i] = q->seqs.size();
        q->seqs.insert(q->seqs.end(), seqs + offsets[i], seqs + offsets[i] + lens[i]);
    }
    q->bytes = estimate_bytes(n_reads, bases);
    (void)s;
    return q;
}

// Cut a request larger than one step into parts of at most max_step_reads
// reads whose estimates fit the budget, in read order. False if a single
// read does not fit the budget.
bool split_request(const gis_server *s, Request *q) {
    const uint32_t n = (uint32_t)q->lens.size();
    uint32_t i = 0;
    while (i < n) {
        uint64_t bases = 0;
        uint32_t k = 0;
        while (i + k < n && k < s->cfg.max_step_reads) {
            const uint64_t L = q->lens[i + k];
            if (k > 0 && estimate_bytes(k + 1, bases + L) > s->cfg.hbm_budget) break;
            bases += L;
            k++;
        }
        if (estimate_bytes(k, bases) > s->cfg.hbm_budget) return false;
        Request *p = new Request();
        p->parent = q;
        q->parts.push_back(p);
        p->lens.assign(q->lens.begin() + i, q->lens.begin() + i + k);
        p->offs.resize(k);
        const uint64_t at = q->offs[i];
        p->seqs.assign(q->seqs.begin() + at, q->seqs.begin() + at + bases);
        for (uint32_t j = 0; j < k; j++) p->offs[j] = q->offs[i + j] - at;
        p->bytes = estimate_bytes(k, bases);
        i += k;
    }
    q->parts_left = q->parts.size();
    std::vector<char>().swap(q->seqs);
    std::vector<uint64_t>().swap(q->offs);
    return true;
}

// Copy a finished request's results out (the parts of a split request one
// after another, offsets rebased). Called by the request's one waiter.
int copy_out(Request *q, gis_nam_out *out) {
    const uint32_t n = (uint32_t)q->lens.size();
    std::vector<Request *> self{q};
    const std::vector<Request *> &pieces = q->parts.empty() ? self : q->parts;
    uint64_t total = 0;
    for (const Request *p : pieces) {
        if (p->status != GIS_OK) return p->status;
        total += p->nams.size();
    }
    if (!out) return GIS_OK;
    out->n_nams = total;
    if (out->nam_capacity < total || (total && !out->nams)) return GIS_ERR_CAPACITY;
    uint64_t at = 0;
    uint32_t r = 0;
    for (const Request *p : pieces) {
        const uint32_t k = (uint32_t)p->lens.size();
        if (out->nam_offsets)
            for (uint32_t i = 0; i < k; i++) out->nam_offsets[r + i] = at + p->nam_off[i];
        if (!p->nams.empty()) std::memcpy(out->nams + at, p->nams.data(), sizeof(gis_nam) * p->nams.size());
        if (out->nonrepetitive_fraction && k) std::memcpy(out->nonrepetitive_fraction + r, p->frac.data(), 4ull * k);
        if (out->flags && k) std::memcpy(out->flags + r, p->flags.data(), k);
        at += p->nams.size();
        r += k;
    }
    if (out->nam_offsets) out->nam_offsets[n] = at;
    return GIS_OK;
}

}  // namespace

extern "C" int gis_server_create(gis_ctx *ctx, const gis_server_config *cfg, gis_server **out) {
    if (!out) return GIS_ERR_BAD_PARAMETER;
    *out = nullptr;
    API_BEGIN(ctx)
    if (!ctx->ix.loaded) return fail(ctx, GIS_ERR_STATE, "no index loaded");
    gis_server_config c{};
    if (cfg) {
        c = *cfg;
    } else {
        c.slots = 6;
        c.max_step_reads = 131072;
        c.hbm_budget = 0;
        gis_map_options_default(&c.options);
    }
    if (c.slots < 1 || c.slots > 16) return fail(ctx, GIS_ERR_BAD_PARAMETER, "slots must be in 1..16");
    if (c.max_step_reads < 1 || c.max_step_reads > (1u << 22))
        return fail(ctx, GIS_ERR_BAD_PARAMETER, "max_step_reads must be in 1..4194304");
    if (c.hbm_budget == 0) c.hbm_budget = 32ull << 30;
    if (c.options.mode != GIS_MODE_FUSED) c.options.mode = GIS_MODE_PLAIN;
    gis_server *s = new gis_server();
    s->ctx = ctx;
    s->cfg = c;
    s->slots.resize(c.slots);
    try {
        for (Slot &sl : s->slots) sl.b = batch_new_stream(ctx);
    } catch (...) {
        for (Slot &sl : s->slots)
            if (sl.b) batch_free(sl.b);
        delete s;
        throw;
    }
    for (size_t i = 0; i < s->slots.size(); i++) s->slots[i].th = std::thread(slot_main, s, i);
    *out = s;
    return GIS_OK;
    API_END(ctx)
}

extern "C" void gis_server_destroy(gis_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        // queued requests that never started are dropped
        for (Request *q : s->pending) {
            q->status = GIS_ERR_CANCELLED;
            q->done = true;
            if (q->parent && --q->parent->parts_left == 0) q->parent->done = true;
        }
        s->pending.clear();
    }
    s->cv_work.notify_all();
    for (Slot &sl : s->slots) sl.th.join();
    {
        // every request is done now; callers still waiting take theirs
        // before the requests are freed
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_done.notify_all();
        s->cv_done.wait(lk, [&] { return s->n_waiters == 0; });
    }
    {
        std::lock_guard<std::recursive_mutex> cl(s->ctx->mu);
        (void)hipSetDevice(s->ctx->device);
        for (Slot &sl : s->slots) {
            (void)hipStreamSynchronize(sl.b->stream);
            batch_free(sl.b);
            sl.release();
        }
    }
    for (auto &kv : s->reqs) delete kv.second;
    delete s;
}

extern "C" int gis_server_submit(gis_server *s, const char *seqs, const uint64_t *offsets,
                                 const uint32_t *lens, uint32_t n_reads, uint64_t *id) {
    if (!s || !id || (n_reads && (!seqs || !offsets || !lens))) return GIS_ERR_BAD_PARAMETER;
    for (uint32_t i = 0; i < n_reads; i++)
        if (lens[i] > GIS_MAX_READ_LEN) return GIS_ERR_BAD_PARAMETER;
    Request *q = nullptr;
    try {
        q = make_request(s, seqs, offsets, lens, n_reads);
    } catch (const std::bad_alloc &) {
        return GIS_ERR_NOMEM;
    }
    if (q->bytes > s->cfg.hbm_budget || n_reads > s->cfg.max_step_reads) {
        bool ok = false;
        try {
            ok = split_request(s, q);
        } catch (const std::bad_alloc &) {
            delete q;
            return GIS_ERR_NOMEM;
        }
        if (!ok) {
            delete q;
            return GIS_ERR_BUDGET;
        }
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) {
            delete q;
            return GIS_ERR_CANCELLED;
        }
        q->id = s->next_id++;
        q->t_submit = Clock::now();
        s->reqs[q->id] = q;
        if (q->parts.empty()) {
            s->pending.push_back(q);
        } else {
            for (Request *p : q->parts) {
                p->t_submit = q->t_submit;
                s->pending.push_back(p);
            }
        }
        s->stats.requests++;
        s->stats.reads += n_reads;
        *id = q->id;
    }
    s->cv_work.notify_all();
    return GIS_OK;
}

extern "C" int gis_server_wait(gis_server *s, uint64_t id, int timeout_ms, gis_nam_out *out,
                               double *latency_ms) {
    if (!s) return GIS_ERR_BAD_PARAMETER;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return GIS_ERR_BAD_PARAMETER;
    Request *q = it->second;
    if (q->claimed) return GIS_ERR_STATE;  // another caller is waiting for it
    q->claimed = true;
    s->n_waiters++;
    auto ready = [&] { return q->done; };
    bool ok = true;
    if (timeout_ms < 0) s->cv_done.wait(lk, ready);
    else ok = s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
    int rc = GIS_ERR_STATE;
    if (ok) {
        // done and claimed: nothing else touches the request now
        lk.unlock();
        rc = copy_out(q, out);
        if (latency_ms) {
            Clock::time_point t_done = q->t_done;
            for (const Request *p : q->parts) t_done = std::max(t_done, p->t_done);
            *latency_ms = std::chrono::duration<double, std::milli>(t_done - q->t_submit).count();
        }
        lk.lock();
    }
    if (ok && rc != GIS_ERR_CAPACITY) {
        s->reqs.erase(id);
        delete q;
    } else {
        q->claimed = false;  // timed out, or the caller retries with more room
    }
    s->n_waiters--;
    s->cv_done.notify_all();
    return rc;
}

extern "C" int gis_server_replay(gis_server *s, const char *seqs, const uint64_t *offsets,
                                 const uint32_t *lens, uint32_t n_requests, const double *arrival_s,
                                 const uint32_t *first, const uint32_t *count, double *latency_ms,
                                 double *wall_s) {
    if (!s || (n_requests && (!seqs || !offsets || !lens || !arrival_s || !first || !count)))
        return GIS_ERR_BAD_PARAMETER;
    std::vector<uint64_t> ids(n_requests, 0);
    std::vector<int> sub_rc(n_requests, GIS_OK);
    std::mutex m;
    std::condition_variable cv;
    uint32_t submitted = 0;
    const Clock::time_point t0 = Clock::now();
    // the client: submits each request at its arrival time
    std::thread client([&] {
        for (uint32_t i = 0; i < n_requests; i++) {
            const Clock::time_point due = t0 + std::chrono::duration_cast<Clock::duration>(
                                                   std::chrono::duration<double>(arrival_s[i]));
            std::this_thread::sleep_until(due);
            uint64_t id = 0;
            const int r = gis_server_submit(s, seqs, offsets + first[i], lens + first[i], count[i], &id);
            {
                std::lock_guard<std::mutex> lk(m);
                ids[i] = id;
                sub_rc[i] = r;
                submitted = i + 1;
            }
            cv.notify_one();
        }
    });
    // this thread collects the requests in arrival order while the client
    // runs, so finished results do not pile up
    int rc = GIS_OK;
    for (uint32_t i = 0; i < n_requests; i++) {
        {
            std::unique_lock<std::mutex> lk(m);
            cv.wait(lk, [&] { return submitted > i; });
        }
        if (sub_rc[i] != GIS_OK) {
            if (rc == GIS_OK) rc = sub_rc[i];
            if (latency_ms) latency_ms[i] = -1.0;
            continue;
        }
        double lat = 0;
        const int w = gis_server_wait(s, ids[i], -1, nullptr, &lat);
        if (w != GIS_OK && rc == GIS_OK) rc = w;
        if (latency_ms) latency_ms[i] = w == GIS_OK ? lat : -1.0;
    }
    client.join();
    if (wall_s) *wall_s = std::chrono::duration<double>(Clock::now() - t0).count();
    return rc;
}

extern "C" int gis_server_stats_get(gis_server *s, gis_server_stats *out) {
    if (!s || !out) return GIS_ERR_BAD_PARAMETER;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    return GIS_OK;
}
