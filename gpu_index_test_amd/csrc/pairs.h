// NAM consumer of paired reads: the reference order of each NAM list, pair
// classification, MAPQ and the joint NAM pairs (step2/main.cu:870-910,
// 1719-1830, 3449-3472, 3552-3581, 4641-4646).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gis.h"

namespace gis {

// largest max_tries: the per-lane "NAM already paired" flags hold 256 bits
constexpr int PAIR_MAX_TRIES = 256;

// Device view of one pair step. Mate 1 of pair p is list p, mate 2 is list
// n_pairs + p of the 2 n_pairs NAM lists at off/in.
struct PairDev {
    const uint64_t *off;  // [2n+1]
    const gis_nam *in;    // input lists
    gis_nam *ord;         // lists in the reference order
    uint32_t n_pairs;
    int grouped;          // input lists sorted by (is_rc, ref_id): rank within a group range only
    uint8_t *types;       // [n]
    float *drop;          // [2n] top_dropoff
    uint8_t *mapq;        // [2n]
    uint32_t *tcnt;       // [5 * nblk] per block type counts, type-major
    const uint64_t *tpos; // [5 * nblk + 1] their exclusive scan
    uint32_t *tids;       // [n] pair ids grouped by type
    uint64_t *ttot;       // [6] type starts in tids and the total
    uint32_t nblk;
    uint32_t *jbound;     // [n] upper bound of a pair's joint entries
    const uint64_t *jboff;// [n+1]
    gis_nam_pair *jtmp;   // entries at jboff
    uint32_t *jcnt;       // [n]
    const uint64_t *joff; // [n+1]
    gis_nam_pair *jout;   // entries at joff
    unsigned *err;        // PAIR_ERR_* bits
    float dropoff, mu, sigma;
    int max_tries;
    int joint;
};

enum : unsigned { PAIR_ERR_DUPLICATE = 1, PAIR_ERR_JOINT = 2 };

constexpr unsigned PAIR_BLOCK = 256;

size_t pair_blocks(uint32_t n_pairs);
// reference order, types, dropoff, MAPQ, type lists and the joint-entry
// bounds (and their scan into jboff)
void launch_pairs_classify(const PairDev &P, uint64_t *scan_tmp, hipStream_t st);
// joint NAM pairs into jtmp, their counts scanned into joff, compacted into jout
void launch_pairs_joint(const PairDev &P, uint64_t *scan_tmp, hipStream_t st);

}  // namespace gis
