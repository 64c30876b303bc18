// Strobemer index construction and the .sti v2 file format.
//
// Construction follows StrobemerIndex::populate (step2/index.cu:143-244):
// randstrobes of every contig (packed = ref_index << 8 | strobe2 offset),
// sorted by (hash, position) -- ties broken by packed so the order is
// deterministic -- then bucket starts over the top `bits` bits and the
// repetitive-seed filter cutoff from the count histogram.
#include "index_host.h"

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

// forward randstrobes of one contig (the reference's RandstrobeGenerator,
// step2/randstrobes.cu:161-190, over the contig's syncmers)
void contig_randstrobes(const char *seq, uint64_t len, uint32_t ref_index, const gis_params &gp,
                        const SeedParams &P, std::vector<gis_ref_randstrobe> &out) {
    std::vector<uint64_t> sy_hash;
    std::vector<uint32_t> sy_pos;
    sy_hash.reserve(len / 4 + 16);
    sy_pos.reserve(len / 4 + 16);
    Ring ring;
    const int k = P.k, s = P.s, w = P.w;
    uint64_t run = 0;  // ACGT bases since the last non-ACGT
    uint64_t fs = 0, rs = 0, fk = 0, rk = 0;
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
        fk = ((fk << 2) | (uint64_t)c) & P.kmask;
        rk = (rk >> 2) | ((uint64_t)(3 - c) << (2 * (k - 1)));
        if (run < (uint64_t)s) continue;
        const int64_t j = (int64_t)(i + 1 - (uint64_t)s);
        ring.v[j & 63] = xxh64_u64(fs < rs ? fs : rs);
        if (run < (uint64_t)k) continue;
        const int64_t p = (int64_t)(i + 1 - (uint64_t)k);
        // window p holds s-mers p .. p+w-1; the templates index h[p + o]
        struct Shift {
            const Ring *r;
            int64_t base;
            uint64_t operator[](int64_t i) const { return (*r)[base + i]; }
        } view{&ring, p};
        if (run == (uint64_t)k) state = anchor_state(view, 0, w, true);
        else state = slide(state, Shift{&ring, p - 1}, 1, w);
        if (state == (uint8_t)(P.t - 1)) {
            sy_hash.push_back(xxh64_u64(fk < rk ? fk : rk));
            sy_pos.push_back((uint32_t)p);
        }
    }
    const size_t n = sy_hash.size();
    const size_t n_out = n > gp.w_min ? n - gp.w_min : 0;
    const size_t o0 = out.size();
    out.resize(o0 + n_out);
    if (len < gp.w_max) {
        // assign_randstrobes returns before writing (step2/index.cu:282-284)
        // while count_randstrobes still counted these slots: they stay zero
        for (size_t i = 0; i < n_out; i++) out[o0 + i] = gis_ref_randstrobe{0, 0, 0};
        return;
    }
    for (size_t i = 0; i < n_out; i++) {
        const size_t wend = std::min(i + (size_t)gp.w_max, n - 1);
        const uint64_t maxp = (uint64_t)sy_pos[i] + P.max_dist_u;
        size_t best = i;
        uint64_t minv = UINT64_MAX;
        for (size_t j = i + gp.w_min; j <= wend && (uint64_t)sy_pos[j] <= maxp; j++) {
            const uint64_t cnt = (uint64_t)__builtin_popcountll((sy_hash[i] ^ sy_hash[j]) & gp.q);
            if (cnt < minv) {
                minv = cnt;
                best = j;
            }
        }
        gis_ref_randstrobe e;
        e.hash = sy_hash[i] + sy_hash[best];
        e.position = sy_pos[i];
        e.packed = (ref_index << 8) + (sy_pos[best] - sy_pos[i]);
        out[o0 + i] = e;
    }
}

inline bool rs_less(const gis_ref_randstrobe &a, const gis_ref_randstrobe &b) {
    if (a.hash != b.hash) return a.hash < b.hash;
    if (a.position != b.position) return a.position < b.position;
    return a.packed < b.packed;
}

void parallel_sort(std::vector<gis_ref_randstrobe> &v, int n_threads) {
    const size_t n = v.size();
    int T = std::max(1, n_threads);
    while (T > 1 && n / (size_t)T < (1u << 16)) T >>= 1;
    if (T <= 1) {
        std::sort(v.begin(), v.end(), rs_less);
        return;
    }
    std::vector<size_t> bounds(T + 1);
    for (int t = 0; t <= T; t++) bounds[t] = n * (size_t)t / (size_t)T;
    std::vector<std::thread> th;
    for (int t = 0; t < T; t++)
        th.emplace_back([&, t]() { std::sort(v.begin() + bounds[t], v.begin() + bounds[t + 1], rs_less); });
    for (auto &x : th) x.join();
    std::vector<gis_ref_randstrobe> tmp(n);
    auto *src = &v, *dst = &tmp;
    for (int width = 1; width < T; width *= 2) {
        th.clear();
        for (int t = 0; t < T; t += 2 * width) {
            const size_t lo = bounds[t], mid = bounds[std::min(T, t + width)],
                         hi = bounds[std::min(T, t + 2 * width)];
            th.emplace_back([=]() {
                std::merge(src->begin() + lo, src->begin() + mid, src->begin() + mid,
                           src->begin() + hi, dst->begin() + lo, rs_less);
            });
        }
        for (auto &x : th) x.join();
        std::swap(src, dst);
    }
    if (src != &v) v.swap(tmp);
}

}  // namespace

void finish_index(HostIndex &ix, float f) {
    const auto &rs = ix.rs;
    const uint64_t n = rs.size();
    const uint64_t nb = (1ULL << ix.bits) + 1;
    ix.starts.clear();
    ix.starts.reserve(nb);
    std::vector<uint64_t> counts;
    ix.tot_occur_once = ix.tot_high_ab = ix.tot_mid_ab = 0;
    uint64_t unique = n == 0 ? 0 : 1;
    uint64_t prev = n == 0 ? 0 : rs[0].hash;
    uint32_t run = 0;
    auto close_run = [&](uint32_t c) {
        if (c == 1) {
            ix.tot_occur_once++;
        } else {
            (c > 100 ? ix.tot_high_ab : ix.tot_mid_ab)++;
            counts.push_back(c);
        }
    };
    // A bucket start is recorded only when the hash changes, so the buckets up
    // to the second distinct hash all start after the first hash's run
    // (step2/index.cu:187-212); lookups reproduce that.
    for (uint64_t i = 0; i < n; i++) {
        const uint64_t h = rs[i].hash;
        if (h == prev) {
            run++;
            continue;
        }
        unique++;
        close_run(run);
        run = 1;
        const uint64_t top = h >> (64 - ix.bits);
        while (ix.starts.size() <= top) ix.starts.push_back(i);
        prev = h;
    }
    close_run(run);
    while (ix.starts.size() < nb) ix.starts.push_back(n);
    std::sort(counts.begin(), counts.end(), std::greater<uint64_t>());
    // index_cutoff = unique_mers * f in float (step2/index.cu:232)
    const uint64_t index_cutoff = (uint64_t)((float)unique * f);
    uint32_t cutoff = 30;
    if (!counts.empty()) {
        cutoff = (uint32_t)(index_cutoff < counts.size() ? counts[index_cutoff] : counts.back());
        cutoff = std::min(100u, std::max(30u, cutoff));
    }
    ix.filter_cutoff = cutoff;
    ix.distinct_strobemers = unique;
    ix.index_cutoff = index_cutoff;
}

int build_host_index(const char *const *seqs, const uint64_t *lens, uint32_t n_refs,
                     const gis_params &p, int bits, float f, int n_threads, HostIndex &out,
                     std::string &err) {
    uint64_t total = 0;
    for (uint32_t j = 0; j < n_refs; j++) total += lens[j];
    if (bits == -1) bits = pick_bits(total, p);
    if (bits < 8 || bits > 31) {
        err = "Bits must be between 8 and 31";
        return GIS_ERR_BAD_PARAMETER;
    }
    if (n_refs >= (1u << 24)) {
        err = "too many reference sequences for the 24-bit ref index";
        return GIS_ERR_BAD_PARAMETER;
    }
    const SeedParams P = make_seed_params(p);
    const int T = std::max(1, n_threads);
    std::vector<std::vector<gis_ref_randstrobe>> per(n_refs);
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < T; t++) {
        th.emplace_back([&]() {
            for (;;) {
                const uint32_t j = next.fetch_add(1);
                if (j >= n_refs) break;
                contig_randstrobes(seqs[j], lens[j], j, p, P, per[j]);
            }
        });
    }
    for (auto &x : th) x.join();
    uint64_t n = 0;
    for (auto &v : per) n += v.size();
    out = HostIndex();
    out.rs.resize(n);
    uint64_t o = 0;
    for (auto &v : per) {
        std::copy(v.begin(), v.end(), out.rs.begin() + (int64_t)o);
        o += v.size();
        std::vector<gis_ref_randstrobe>().swap(v);
    }
    parallel_sort(out.rs, T);
    out.bits = bits;
    out.params = p;
    finish_index(out, f);
    return GIS_OK;
}

bool params_equal(const gis_params &a, const gis_params &b) {
    return a.canonical_read_length == b.canonical_read_length && a.k == b.k && a.s == b.s &&
           a.t_syncmer == b.t_syncmer && a.l == b.l && a.u == b.u && a.q == b.q &&
           a.max_dist == b.max_dist && a.w_min == b.w_min && a.w_max == b.w_max;
}

// .sti v2: "STI\1", u32 version, u64 reserved size + bytes, i32 filter_cutoff,
// i32 bits, 7 x i32 parameters, u64 n + n x 16 B randstrobes, u64 m + m x 8 B
// bucket starts (step2/index.cu:73-134, step2/indexparameters.cpp:85-104)
int write_sti(const std::string &path, const HostIndex &ix, std::string &err) {
    FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) {
        err = path + ": " + std::strerror(errno);
        return GIS_ERR_IO;
    }
    auto w32 = [&](int32_t v) { std::fwrite(&v, 4, 1, fp); };
    auto w64 = [&](uint64_t v) { std::fwrite(&v, 8, 1, fp); };
    std::fwrite("STI\1", 1, 4, fp);
    w32(2);
    w64(8);
    const char zeros[8] = {0};
    std::fwrite(zeros, 1, 8, fp);
    w32((int32_t)ix.filter_cutoff);
    w32(ix.bits);
    const gis_params &p = ix.params;
    w32(p.canonical_read_length);
    w32(p.k);
    w32(p.s);
    w32(p.l);
    w32(p.u);
    w32((int32_t)p.q);
    w32(p.max_dist);
    w64(ix.rs.size());
    if (!ix.rs.empty()) std::fwrite(ix.rs.data(), sizeof(gis_ref_randstrobe), ix.rs.size(), fp);
    w64(ix.starts.size());
    if (!ix.starts.empty()) std::fwrite(ix.starts.data(), 8, ix.starts.size(), fp);
    const bool bad = std::ferror(fp) != 0;
    std::fclose(fp);
    if (bad) {
        err = path + ": write failed";
        return GIS_ERR_IO;
    }
    return GIS_OK;
}

int read_sti(const std::string &path, const gis_params *expected, HostIndex &out, std::string &err) {
    FILE *fp = std::fopen(path.c_str(), "rb");
    if (!fp) {
        err = path + ": " + std::strerror(errno);
        return GIS_ERR_INVALID_INDEX;
    }
    auto fail = [&](const std::string &m) {
        std::fclose(fp);
        err = m;
        return GIS_ERR_INVALID_INDEX;
    };
    char magic[4];
    int32_t v32[9];
    uint64_t v64;
    if (std::fread(magic, 1, 4, fp) != 4 || std::memcmp(magic, "STI\1", 4) != 0)
        return fail("Index file has incorrect format (magic number mismatch)");
    uint32_t version;
    if (std::fread(&version, 4, 1, fp) != 1) return fail("truncated index file");
    if (version != 2)
        return fail("Can only read index file format version 2, but found version " +
                    std::to_string(version));
    if (std::fread(&v64, 8, 1, fp) != 1 || std::fseek(fp, (long)v64, SEEK_CUR) != 0)
        return fail("truncated index file");
    if (std::fread(v32, 4, 9, fp) != 9) return fail("truncated index file");
    gis_params p;
    if (gis_params_make(v32[2], v32[3], v32[4], v32[5], v32[6], v32[7], v32[8], &p) != GIS_OK)
        return fail("invalid index parameters in .sti file");
    if (expected && !params_equal(p, *expected))
        return fail("Index parameters in .sti file and those specified on command line differ");
    out = HostIndex();
    out.filter_cutoff = (uint32_t)v32[0];
    out.bits = v32[1];
    out.params = p;
    if (out.bits < 8 || out.bits > 31) return fail("invalid bits in .sti file");
    if (std::fread(&v64, 8, 1, fp) != 1) return fail("truncated index file");
    out.rs.resize(v64);
    if (v64 && std::fread(out.rs.data(), sizeof(gis_ref_randstrobe), v64, fp) != v64)
        return fail("truncated index file");
    if (std::fread(&v64, 8, 1, fp) != 1) return fail("truncated index file");
    out.starts.resize(v64);
    if (v64 && std::fread(out.starts.data(), 8, v64, fp) != v64) return fail("truncated index file");
    std::fclose(fp);
    if (out.starts.size() != (1ULL << out.bits) + 1) {
        err = "randstrobe_start_indices vector is of the wrong size";
        return GIS_ERR_INVALID_INDEX;
    }
    return GIS_OK;
}

}  // namespace gis
