// Host side of the strobemer index: builder, .sti v2 I/O.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/gis.h"

namespace gis {

struct HostIndex {
    std::vector<gis_ref_randstrobe> rs;
    std::vector<uint64_t> starts;
    int bits = 0;
    uint32_t filter_cutoff = 0;
    gis_params params{};
    uint64_t distinct_strobemers = 0, tot_occur_once = 0, tot_high_ab = 0, tot_mid_ab = 0,
             index_cutoff = 0;
};

int pick_bits(uint64_t total_length, const gis_params &p);
// repetition statistics and filter_cutoff from the run-length histogram
// (hist[1..100] exact, hist[101] = runs longer than 100)
void filter_stats(const uint64_t *hist, float f, HostIndex &ix);
// a .sti v2 file mapped read-only: the table and starts point into the mapping
struct MappedSti {
    void *base = nullptr;
    uint64_t size = 0;
    const gis_ref_randstrobe *rs = nullptr;
    uint64_t n = 0;
    const uint64_t *starts = nullptr;  // 2^bits + 1 entries (checked non-decreasing, <= n)
    int bits = 0;
    uint32_t filter_cutoff = 0;
    gis_params params{};
};
int map_sti(const std::string &path, const gis_params *expected, MappedSti &out, std::string &err);
void unmap_sti(MappedSti &m);
bool params_equal(const gis_params &a, const gis_params &b);

}  // namespace gis
