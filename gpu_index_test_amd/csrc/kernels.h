// Kernel-side data structures and launch wrappers of the seeding path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gis_core.h"

namespace gis {

// Resident index: RefRandstrobe table (16 B records: u64 hash, u32 position,
// u32 packed) sorted by (hash, position), and bucket starts of 2^bits + 1
// entries (step2/index.hpp:168-182).
struct DevIndex {
    const uint4 *rs;
    const uint64_t *starts;
    uint64_t n;
    int bits;
    uint32_t filter_cutoff;
    int k;
};

// One hit (step2/my_struct.hpp:71-86) with its grouping key:
// ref_strand = is_rc << 31 | ref_id.
struct DevHit {
    uint32_t ref_strand;
    int32_t ref_start;
    int32_t ref_end;
    uint16_t query_start;
    uint16_t query_end;
};

// Per-batch device arrays of the per-stage pipeline.
struct DevBatch {
    // reads
    const char *seqs;
    const uint64_t *read_off;
    const uint32_t *read_len;
    uint32_t n_reads;
    // randstrobes: per-read regions of capacity from the read length
    const uint64_t *rs_off;   // [n+1] capacity offsets
    uint32_t *rs_cnt;         // [n]
    uint32_t *rs_nfwd;        // [n] forward-strand randstrobes (they come first)
    uint64_t *rs_hash;
    uint32_t *rs_start;
    uint32_t *rs_end;
    // optional syncmer output (stage entry point), regions at sy_off
    const uint64_t *sy_off;
    uint32_t *sy_cnt;
    uint64_t *sy_hash;
    uint64_t *sy_pos;
    // lookup results per randstrobe slot
    uint64_t *pos;            // GIS_NOT_FOUND if absent
    uint32_t *lk_bytes;       // [n] algorithmic bytes the lookup kernel moved for the read
    uint32_t *hcnt;           // hits this seed contributes (rescued reads: run length)
    uint32_t *rcnt;           // rescued reads: hits this seed contributes
    // per read
    uint32_t *hit_total;      // normal path hits
    float *frac;
    uint8_t *flags;
    uint32_t *rhit_total;     // rescue hits
    // hits
    const uint64_t *hit_off;  // [n+1]
    DevHit *hits;
    const uint64_t *rhit_off; // [n+1]
    DevHit *rhits;
    // merge scratch (indexed like the hit arrays)
    uint64_t *keys;           // 2 per hit
    uint32_t *aux;            // 2 per hit (group starts / open list)
    gis_nam *nam_ws;          // working NAMs
    gis_nam *nam_fin;         // finished NAMs in canonical order
    gis_nam *rnam_ws;
    gis_nam *rnam_fin;
    uint64_t *rkeys;
    uint32_t *raux;
    uint32_t *nam_cnt;        // per read
    // output
    const uint64_t *nam_off;  // [n+1]
    gis_nam *nams_out;
    // long-read scratch for the seeding stage
    const uint64_t *long_off; // [n+1] per-read scratch offsets in bases (0 for short reads)
    uint8_t *long_scratch;
};

constexpr int LMAX_LDS = 320;  // reads up to this length seed from LDS
constexpr int SEED_WPB = 4;    // waves per block of the seeding kernel
constexpr int WPB = 4;         // waves per block of the per-read kernels
// bytes of long-read scratch per base: codes(1) + state(1) + s-mer hash(8)
// + syncmer hash(8) + syncmer pos(4) + N mask (1/8 per base, rounded up)
constexpr int LONG_BYTES_PER_BASE = 24;

// launch wrappers (kernels.hip)
void launch_seed(const DevBatch &b, const SeedParams &sp, int want_syncmers, int long_reads,
                 hipStream_t st);
// rescue_mode: 0 = the caller's decision (nonrepetitive fraction < 0.7 or no
// hits), 1 = never rescue, 2 = always rescue
void launch_lookup(const DevBatch &b, const DevIndex &ix, int rescue_mode, hipStream_t st);
void launch_hits(const DevBatch &b, const DevIndex &ix, hipStream_t st);
void launch_merge(const DevBatch &b, int rescue, int k, hipStream_t st);
void launch_rescue_count(const DevBatch &b, const DevIndex &ix, uint32_t rescue_cutoff,
                         uint32_t hard_cutoff, hipStream_t st);
void launch_rescue_hits(const DevBatch &b, const DevIndex &ix, uint32_t rescue_cutoff,
                        uint32_t hard_cutoff, hipStream_t st);
void launch_compact(const DevBatch &b, hipStream_t st);
void launch_find_flat(const DevIndex &ix, const uint64_t *hashes, uint64_t n, uint64_t *positions,
                      hipStream_t st);
void launch_rs_capacity(const uint32_t *lens, uint32_t n, const SeedParams &sp, uint32_t *cap_rs,
                        uint32_t *cap_sy, hipStream_t st);
// exclusive scan of u32 counts into u64 offsets[n+1] (offsets[n] = total);
// tmp must hold scan_tmp_words(n) u64.
size_t scan_tmp_words(uint64_t n);
void launch_scan(const uint32_t *in, uint64_t *out, uint64_t n, uint64_t *tmp, hipStream_t st);

}  // namespace gis
