// Device half of strobemer index construction (index_build.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gis.h"

namespace gis {

constexpr int RUN_BINS = 102;  // run-length histogram: [1..100] exact, [101] longer runs

struct BuildError {
    hipError_t e;
    const char *what;
};

// The RefRandstrobe records of every contig (RandstrobeGenerator over each
// contig, step2/randstrobes.cu:161-190, step2/index.cu:246-311), seeded on
// the device in (contig, position) order; contigs shorter than w_max get
// zero records in their slots. *out is a device array of *n_out records
// from hipMalloc, owned by the caller.
void device_seed_reference(const char *const *seqs, const uint64_t *lens, uint32_t n_refs, const gis_params &p,
                           uint4 **out, uint64_t *n_out, hipStream_t st);
// sort n RefRandstrobe records (16 B: u64 hash, u32 position, u32 packed) in
// place by (hash, position, packed); allocates and frees its own buffers
void device_sort_records(uint4 *rs, uint64_t n, hipStream_t st);
// checks a resident table and its starts: null if the lookups can trust
// them, else what is wrong (unsorted records, or bucket starts that do not
// match the records' top hash bits)
const char *device_validate_index(const uint4 *rs, uint64_t n, const uint64_t *starts, int bits, hipStream_t st);
// bucket starts (2^bits + 1 entries) and the run-length histogram of a
// sorted table; hist_out receives RUN_BINS host counters
void device_finish_index(const uint4 *rs, uint64_t n, int bits, uint64_t *starts, uint64_t *hist_out,
                         hipStream_t st);

}  // namespace gis
