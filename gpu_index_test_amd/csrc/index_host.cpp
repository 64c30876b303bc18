// Host side of strobemer index construction, and the .sti v2 file format.
//
// StrobemerIndex::populate (step2/index.cu:143-244) is split between the
// host and the GPU: the host seeds the reference (randstrobes of every
// contig, packed = ref_index << 8 | strobe2 offset) in parallel segments;
// the device sorts the records by (hash, position, packed), fills the
// bucket starts and counts runs of equal hashes (index_build.hip); the host
// turns that run-length histogram into the repetition statistics and the
// filter cutoff.
#include "index_host.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <thread>

#include "gis_core.h"

namespace gis {

int pick_bits(uint64_t total_length, const gis_params &p) {
    // two randstrobes per bucket on average, step2/index.cu:137-141
    const uint64_t est = total_length / (uint64_t)(p.k - p.s + 1);
    if (est == 0) return 8;
    const int b = (int)std::log2((double)est) - 1;
    return std::min(31, std::max(8, b));
}

namespace {

// ring view of the last 64 s-mer hashes, indexed by absolute s-mer start
struct Ring {
    uint64_t v[64];
    uint64_t operator[](int64_t i) const { return v[i & 63]; }
};

// Randstrobes (RefRandstrobe, packed = ref_index << 8 | strobe2 offset) of
// one contig whose strobe1 syncmer starts in [A, B): the reference's
// RandstrobeGenerator (step2/randstrobes.cu:161-190) over the contig's
// syncmers, restricted to a segment so that one contig can be seeded by
// several threads. The syncmer scan starts `back` bases before A and is
// exact once it has passed a window whose state needs no history (a unique
// minimum below ~0, or the first window after a contig start or an N);
// returns false when no such window lies before A, and the caller retries
// with a longer run-in. The scan continues past B until the partner windows
// of the last strobe1 are complete (max_dist) and w_min syncmers follow it.
bool segment_randstrobes(const char *seq, uint64_t len, uint32_t ref_index, uint64_t A, uint64_t B,
                         uint64_t back, const gis_params &gp, const SeedParams &P,
                         std::vector<gis_ref_randstrobe> &out) {
    const int k = P.k, s = P.s, w = P.w;
    const uint64_t S = A > back ? A - back : 0;
    std::vector<uint64_t> sy_hash;
    std::vector<uint32_t> sy_pos;
    uint64_t first_owned = UINT64_MAX;  // index in sy_* of the first syncmer at >= A
    uint64_t after_B = 0;
    Ring ring;
    uint64_t run = 0;  // ACGT bases since the last non-ACGT or the scan start
    uint64_t fs = 0, rs = 0, fk = 0, rk = 0;
    uint8_t state = ST_INVALID;
    bool exact = S == 0;  // at the contig start the first window is a true segment start
    const uint64_t stop_pos = B + P.max_dist_u + 1;
    for (uint64_t i = S; i < len; i++) {
        const int c = nt4((unsigned char)seq[i]);
        if (c > 3) {
            run = 0;
            state = ST_INVALID;
            exact = true;  // the next window starts a segment
            continue;
        }
        run++;
        fs = ((fs << 2) | (uint64_t)c) & P.smask;
        rs = (rs >> 2) | ((uint64_t)(3 - c) << (2 * (s - 1)));
        fk = ((fk << 2) | (uint64_t)c) & P.kmask;
        rk = (rk >> 2) | ((uint64_t)(3 - c) << (2 * (k - 1)));
        if (run < (uint64_t)s) continue;
        const int64_t j = (int64_t)(i + 1 - (uint64_t)s);
        ring.v[j & 63] = xxh64_u64(fs < rs ? fs : rs);
        if (run < (uint64_t)k) continue;
        const uint64_t p = i + 1 - (uint64_t)k;
        struct Shift {
            const Ring *r;
            int64_t base;
            uint64_t operator[](int64_t x) const { return (*r)[base + x]; }
        };
        if (run == (uint64_t)k) {
            state = anchor_state(Shift{&ring, (int64_t)p}, 0, w, true);
        } else {
            state = slide(state, Shift{&ring, (int64_t)p - 1}, 1, w);
            if (!exact && anchor_state(Shift{&ring, (int64_t)p}, 0, w, false) != ST_UNKNOWN) exact = true;
        }
        if (p == A && !exact) return false;
        if (p >= A && !exact) return false;  // (A itself may be inside an N run)
        if (state == (uint8_t)(P.t - 1)) {
            if (p >= A && first_owned == UINT64_MAX) first_owned = sy_hash.size();
            if (p >= B) after_B++;
            sy_hash.push_back(xxh64_u64(fk < rk ? fk : rk));
            sy_pos.push_back((uint32_t)p);
        }
        if (p >= stop_pos && after_B >= gp.w_min) break;
    }
    if (first_owned == UINT64_MAX) return true;
    const size_t n = sy_hash.size();
    for (size_t i = first_owned; i < n && sy_pos[i] < B; i++) {
        if (i + gp.w_min > n - 1) break;  // fewer than w_min syncmers follow: no randstrobe
        const size_t wend = std::min(i + (size_t)gp.w_max, n - 1);
        const uint64_t maxp = (uint64_t)sy_pos[i] + P.max_dist_u;
        size_t best = i;
        uint64_t minv = UINT64_MAX;
        for (size_t jj = i + gp.w_min; jj <= wend && (uint64_t)sy_pos[jj] <= maxp; jj++) {
            const uint64_t cnt = (uint64_t)__builtin_popcountll((sy_hash[i] ^ sy_hash[jj]) & gp.q);
            if (cnt < minv) {
                minv = cnt;
                best = jj;
            }
        }
        gis_ref_randstrobe e;
        e.hash = sy_hash[i] + sy_hash[best];
        e.position = sy_pos[i];
        e.packed = (ref_index << 8) + (sy_pos[best] - sy_pos[i]);
        out.push_back(e);
    }
    return true;
}

// count of randstrobe slots of a contig shorter than w_max: the reference
// counts them (count_randstrobes) but assign_randstrobes returns before
// writing (step2/index.cu:282-284), so they stay zero
uint64_t short_contig_slots(const char *seq, uint64_t len, const gis_params &gp, const SeedParams &P) {
    // count syncmers with a full scan (the contig is tiny)
    uint64_t nsy = 0;
    {
        const int k = P.k, s = P.s, w = P.w;
        Ring ring;
        uint64_t run = 0, fs = 0, rs = 0;
        uint8_t state = ST_INVALID;
        for (uint64_t i = 0; i < len; i++) {
            const int c = nt4((unsigned char)seq[i]);
            if (c > 3) {
                run = 0;
                state = ST_INVALID;
                continue;
            }
            run++;
            fs = ((fs << 2) | (uint64_t)c) & P.smask;
            rs = (rs >> 2) | ((uint64_t)(3 - c) << (2 * (s - 1)));
            if (run < (uint64_t)s) continue;
            const int64_t j = (int64_t)(i + 1 - (uint64_t)s);
            ring.v[j & 63] = xxh64_u64(fs < rs ? fs : rs);
            if (run < (uint64_t)k) continue;
            const int64_t p = (int64_t)(i + 1 - (uint64_t)k);
            struct Shift {
                const Ring *r;
                int64_t base;
                uint64_t operator[](int64_t x) const { return (*r)[base + x]; }
            };
            if (run == (uint64_t)k) state = anchor_state(Shift{&ring, p}, 0, w, true);
            else state = slide(state, Shift{&ring, p - 1}, 1, w);
            if (state == (uint8_t)(P.t - 1)) nsy++;
        }
    }
    return nsy > gp.w_min ? nsy - gp.w_min : 0;
}

}  // namespace

int seed_reference(const char *const *seqs, const uint64_t *lens, uint32_t n_refs, const gis_params &p,
                   int n_threads, uint64_t segment, std::vector<gis_ref_randstrobe> &out, std::string &err) {
    if (n_refs >= (1u << 24)) {
        err = "too many reference sequences for the 24-bit ref index";
        return GIS_ERR_BAD_PARAMETER;
    }
    if (segment == 0) segment = 4000000;
    const SeedParams P = make_seed_params(p);
    // jobs: [ref, A, B) segments in (contig, position) order
    struct Job {
        uint32_t ref;
        uint64_t A, B;
    };
    std::vector<Job> jobs;
    for (uint32_t j = 0; j < n_refs; j++) {
        if (lens[j] < p.w_max) {
            jobs.push_back(Job{j, 0, 0});  // zero slots of a contig shorter than w_max
            continue;
        }
        for (uint64_t a = 0; a < lens[j]; a += segment) jobs.push_back(Job{j, a, std::min(lens[j], a + segment)});
    }
    std::vector<std::vector<gis_ref_randstrobe>> per(jobs.size());
    std::atomic<size_t> next{0};
    std::vector<std::thread> th;
    const int T = std::max(1, std::min<int>(n_threads, (int)jobs.size()));
    for (int t = 0; t < T; t++) {
        th.emplace_back([&]() {
            for (;;) {
                const size_t x = next.fetch_add(1);
                if (x >= jobs.size()) break;
                const Job &jb = jobs[x];
                auto &o = per[x];
                if (jb.B == 0) {
                    o.assign(short_contig_slots(seqs[jb.ref], lens[jb.ref], p, P), gis_ref_randstrobe{0, 0, 0});
                    continue;
                }
                uint64_t back = 1024;
                while (!segment_randstrobes(seqs[jb.ref], lens[jb.ref], jb.ref, jb.A, jb.B, back, p, P, o)) {
                    o.clear();
                    back *= 4;
                }
            }
        });
    }
    for (auto &x : th) x.join();
    uint64_t n = 0;
    for (auto &v : per) n += v.size();
    out.resize(n);
    uint64_t o = 0;
    for (auto &v : per) {
        std::copy(v.begin(), v.end(), out.begin() + (int64_t)o);
        o += v.size();
        std::vector<gis_ref_randstrobe>().swap(v);
    }
    return GIS_OK;
}

void filter_stats(const uint64_t *hist, float f, HostIndex &ix) {
    // hist[c] = runs of c equal hashes for c = 1..100, hist[101] = runs of
    // more than 100 (step2/index.cu:214-244: counts of the repeated hashes
    // sorted in descending order, index_cutoff = unique * f in float, the
    // count at that rank clamped to [30, 100]).
    uint64_t unique = 0, repeated = 0;
    for (int c = 1; c <= 101; c++) unique += hist[c];
    for (int c = 2; c <= 101; c++) repeated += hist[c];
    ix.tot_occur_once = hist[1];
    ix.tot_high_ab = hist[101];
    ix.tot_mid_ab = repeated - hist[101];
    ix.distinct_strobemers = unique;
    const uint64_t index_cutoff = (uint64_t)((float)unique * f);
    ix.index_cutoff = index_cutoff;
    uint32_t cutoff = 30;
    if (repeated > 0) {
        uint64_t v = 0;
        if (index_cutoff < repeated) {
            if (index_cutoff < hist[101]) {
                v = 101;
            } else {
                uint64_t cum = hist[101];
                for (int c = 100; c >= 2; c--) {
                    cum += hist[c];
                    if (index_cutoff < cum) {
                        v = (uint64_t)c;
                        break;
                    }
                }
            }
        } else {
            v = 101;  // the smallest repeated count
            for (int c = 2; c <= 100; c++)
                if (hist[c]) {
                    v = (uint64_t)c;
                    break;
                }
        }
        cutoff = (uint32_t)std::min<uint64_t>(100, std::max<uint64_t>(30, v));
    }
    ix.filter_cutoff = cutoff;
}

bool params_equal(const gis_params &a, const gis_params &b) {
    return a.canonical_read_length == b.canonical_read_length && a.k == b.k && a.s == b.s &&
           a.t_syncmer == b.t_syncmer && a.l == b.l && a.u == b.u && a.q == b.q &&
           a.max_dist == b.max_dist && a.w_min == b.w_min && a.w_max == b.w_max;
}

// .sti v2: "STI\1", u32 version, u64 reserved size + bytes, i32 filter_cutoff,
// i32 bits, 7 x i32 parameters, u64 n + n x 16 B randstrobes, u64 m + m x 8 B
// bucket starts (step2/index.cu:73-134, step2/indexparameters.cpp:85-104)
// The .sti v2 file mapped into memory: header fields parsed and checked, the
// table and the bucket starts left in the mapping (no host copy).
int map_sti(const std::string &path, const gis_params *expected, MappedSti &out, std::string &err) {
    out = MappedSti();
    const int fd = ::open(path.c_str(), O_RDONLY);
    if (fd < 0) {
        err = path + ": " + std::strerror(errno);
        return GIS_ERR_INVALID_INDEX;
    }
    struct stat st;
    if (::fstat(fd, &st) != 0) {
        err = path + ": " + std::strerror(errno);
        ::close(fd);
        return GIS_ERR_INVALID_INDEX;
    }
    const uint64_t size = (uint64_t)st.st_size;
    void *base = size ? ::mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0) : MAP_FAILED;
    ::close(fd);
    if (base == MAP_FAILED) {
        err = "Index file has incorrect format (magic number mismatch)";
        return GIS_ERR_INVALID_INDEX;
    }
    ::madvise(base, size, MADV_SEQUENTIAL);
    out.base = base;
    out.size = size;
    const unsigned char *p = static_cast<const unsigned char *>(base);
    uint64_t at = 0;
    auto take = [&](void *dst, uint64_t n) {
        if (at + n > size) return false;
        std::memcpy(dst, p + at, n);
        at += n;
        return true;
    };
    auto fail = [&](const std::string &m) {
        unmap_sti(out);
        err = m;
        return GIS_ERR_INVALID_INDEX;
    };
    char magic[4];
    if (!take(magic, 4) || std::memcmp(magic, "STI\1", 4) != 0)
        return fail("Index file has incorrect format (magic number mismatch)");
    uint32_t version;
    if (!take(&version, 4)) return fail("truncated index file");
    if (version != 2)
        return fail("Can only read index file format version 2, but found version " + std::to_string(version));
    uint64_t v64;
    if (!take(&v64, 8) || at + v64 > size) return fail("truncated index file");
    at += v64;  // reserved chunk
    int32_t v32[9];
    if (!take(v32, sizeof v32)) return fail("truncated index file");
    gis_params prm;
    if (gis_params_make(v32[2], v32[3], v32[4], v32[5], v32[6], v32[7], v32[8], &prm) != GIS_OK)
        return fail("invalid index parameters in .sti file");
    if (expected && !params_equal(prm, *expected))
        return fail("Index parameters in .sti file and those specified on command line differ");
    out.filter_cutoff = (uint32_t)v32[0];
    out.bits = v32[1];
    out.params = prm;
    if (out.bits < 8 || out.bits > 31) return fail("invalid bits in .sti file");
    if (!take(&v64, 8) || v64 > (size - at) / sizeof(gis_ref_randstrobe)) return fail("truncated index file");
    out.n = v64;
    out.rs = reinterpret_cast<const gis_ref_randstrobe *>(p + at);
    at += v64 * sizeof(gis_ref_randstrobe);
    if (!take(&v64, 8) || v64 > (size - at) / 8) return fail("truncated index file");
    if (v64 != (1ULL << out.bits) + 1) return fail("randstrobe_start_indices vector is of the wrong size");
    out.starts = reinterpret_cast<const uint64_t *>(p + at);
    // the lookups index the table through the bucket starts: they must be
    // non-decreasing and within the table
    uint64_t prev = 0;
    for (uint64_t i = 0; i < v64; i++) {
        uint64_t v;
        std::memcpy(&v, p + at + 8 * i, 8);
        if (v < prev || v > out.n) return fail("bucket starts out of range");
        prev = v;
    }
    return GIS_OK;
}

void unmap_sti(MappedSti &m) {
    if (m.base) ::munmap(m.base, m.size);
    m = MappedSti();
}

}  // namespace gis
