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
int build_host_index(const char *const *seqs, const uint64_t *lens, uint32_t n_refs,
                     const gis_params &p, int bits, float f, int n_threads, HostIndex &out,
                     std::string &err);
// bucket starts, repetition statistics and filter_cutoff of a sorted table
void finish_index(HostIndex &ix, float f);
int read_sti(const std::string &path, const gis_params *expected, HostIndex &out, std::string &err);
int write_sti(const std::string &path, const HostIndex &ix, std::string &err);
bool params_equal(const gis_params &a, const gis_params &b);

}  // namespace gis
