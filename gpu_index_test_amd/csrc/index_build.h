// Device half of strobemer index construction (index_build.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gis {

constexpr int RUN_BINS = 102;  // run-length histogram: [1..100] exact, [101] longer runs

struct BuildError {
    hipError_t e;
    const char *what;
};

// sort n RefRandstrobe records (16 B: u64 hash, u32 position, u32 packed) in
// place by (hash, position, packed); allocates and frees its own buffers
void device_sort_records(uint4 *rs, uint64_t n, hipStream_t st);
// bucket starts (2^bits + 1 entries) and the run-length histogram of a
// sorted table; hist_out receives RUN_BINS host counters
void device_finish_index(const uint4 *rs, uint64_t n, int bits, uint64_t *starts, uint64_t *hist_out,
                         hipStream_t st);

}  // namespace gis
