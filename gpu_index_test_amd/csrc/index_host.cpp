// Host side of strobemer index construction, and the .sti v2 file format.
//
// StrobemerIndex::populate (step2/index.cu:143-244) runs on the device:
// the reference is seeded there (ref_seed.hip), the records are sorted by
// (hash, position, packed), the bucket starts filled and the runs of equal
// hashes counted (index_build.hip). The host picks the bucket bits and turns
// the run-length histogram into the repetition statistics and the filter
// cutoff.
#include "index_host.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>

#include "gis_core.h"

namespace gis {

int pick_bits(uint64_t total_length, const gis_params &p) {
    // two randstrobes per bucket on average, step2/index.cu:137-141
    const uint64_t est = total_length / (uint64_t)(p.k - p.s + 1);
    if (est == 0) return 8;
    const int b = (int)std::log2((double)est) - 1;
    return std::min(31, std::max(8, b));
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
    if (!take(&v64, 8) || v64 > size - at) return fail("truncated index file");
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
