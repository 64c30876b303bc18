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
    // bucket lines: 2^bits lines of BUCKET_LINE uint4 (128 bytes, one cache
    // line each): {start lo, start hi, size, 0} and the bucket's first
    // BUCKET_LINE - 1 records (zeros past its size). The lookup reads a
    // bucket's bounds and records with one line fetch.
    const uint4 *lines;
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

// NAM under construction on the merge's global list (groups with more than
// 64 open NAMs): query start/end/previous hit start, ref start/end/previous
// hit start, hits so far, creation index within the group
struct WideNam {
    int32_t qs, qe, qp, rs, re, rp;
    uint32_t n, cid;
};

// Device control block of one batch run: the spill list length of the fused
// kernel and the overflow bits of the capacity checks. Reset by k_reset_ctl.
struct DevCtl {
    unsigned int n_spill;   // reads handed from the fused kernel to the per-stage kernels
    unsigned int overflow;  // OVF_* bits
    unsigned int q_head[2]; // dynamic work-queue heads of the fused kernel's two bins
    unsigned int n_gwork[2];  // group work items of split reads (normal, rescue merge)
    unsigned int n_split[2];  // split reads (normal, rescue merge)
    unsigned int n_sel[3];    // selected reads: rescued, deferred merges, pooled hits
    unsigned long long hit_pool;  // hits placed in the pool after the per-read regions
    // totals the host reads with this block: end of the hit pool, rescue
    // hits, NAMs
    unsigned long long tot[3];
};

enum : unsigned { OVF_HITS = 1, OVF_RHITS = 2, OVF_MERGE = 4, OVF_OUT = 8 };

// the public GIS_FLAG_SPILLED bit (include/gis.h)
enum : uint8_t { FLAG_SPILL = GIS_FLAG_SPILLED };

// The reads a per-read kernel visits: all of [0, n) (ids == nullptr) or the
// first *n_dev entries of ids.
struct ReadSet {
    const uint32_t *ids;
    const unsigned int *n_dev;
    uint32_t n;
};

// Per-batch device arrays.
struct DevBatch {
    // reads
    const char *seqs;
    const uint64_t *read_off;
    const uint32_t *read_len;
    uint32_t n_reads;
    // randstrobes: per-read regions, capacity from the read length
    const uint64_t *rs_off;
    uint32_t *rs_cnt;
    uint32_t *rs_nfwd;        // forward-strand randstrobes (they come first)
    uint64_t *rs_hash;
    uint32_t *rs_span;        // query start | query end << 16 (read positions fit 16 bits)
    // optional syncmer output (stage entry point), regions at sy_off
    const uint64_t *sy_off;
    uint32_t *sy_cnt;
    uint64_t *sy_hash;
    uint64_t *sy_pos;
    // lookup results per randstrobe slot
    uint64_t *pos;            // GIS_NOT_FOUND if absent
    uint32_t *hcnt;           // hits this seed contributes (rescued reads: run length)
    uint32_t *rcnt;           // rescued reads: hits this seed contributes
    uint32_t *lk_bytes;       // [n] algorithmic table bytes of the read's lookups (byte model of the search)
    uint32_t *lk_stream;      // [n] bytes the lookup kernel streams beside them: queries, results, hits
    // per read
    uint32_t *hit_total;
    uint32_t *rhit_total;
    float *frac;
    uint8_t *flags;
    uint32_t *nam_cnt;
    uint64_t *hit_off;        // [n+1] the read's hits start here: its own region r * HIT_REGION, or
                              // (more hits than that) a place in the pool after the regions;
                              // hit_off[n] = end of the pool
    uint64_t *rhit_off;       // [n+1] scan of rhit_total
    // hit arrays (capacity hit_cap / rhit_cap hits); merge scratch per read r
    // with h hits at hit offset o: keys[2o .. 2o+2h), nam_ws[o .. o+h)
    DevHit *hits;
    DevHit *rhits;
    uint64_t hit_cap, rhit_cap;
    uint64_t *keys, *rkeys;
    WideNam *nam_ws, *rnam_ws;
    uint32_t merge_scratch;   // keys/nam_ws/gwork allocated; else reads needing them set OVF_MERGE
    // reads with many hits merge each (strand, ref) group on its own wave:
    // gwork items {read, first, end, group | order code << 30}, split list
    // entries {read, groups, order code, 0}
    uint4 *gwork;
    uint4 *split_list;
    // ids of the reads of the sparse passes: rescued at [0, n), deferred
    // merges at [n, 2n), reads whose hits go to the pool at [2n, 3n)
    uint32_t *sel_ids;
    gis_nam *nam_fin, *rnam_fin;
    // fused kernel: NAM slots of F_NAMCAP per read
    gis_nam *nam_slots;
    // dense output
    uint64_t *nam_off;        // [n+1]
    gis_nam *nams_out;
    uint64_t out_cap;
    // long-read scratch for the seeding stage (byte offsets)
    const uint64_t *long_off;
    uint8_t *long_scratch;
    DevCtl *ctl;
    uint32_t *spill_ids;
};

// a query randstrobe's start and end in one word (rs_span)
__host__ __device__ inline uint32_t span_pack(uint32_t start, uint32_t end) { return start | (end << 16); }
__host__ __device__ inline uint32_t span_start(uint32_t v) { return v & 0xffffu; }
__host__ __device__ inline uint32_t span_end(uint32_t v) { return v >> 16; }

constexpr int LMAX_LDS = 320;  // reads up to this length seed from LDS
constexpr int SEED_WPB = 4;    // waves per block of the seeding kernel
constexpr int WPB = 4;         // waves per block of the per-read kernels
constexpr int MERGE_LDS = 128; // hits a read may have for the LDS merge
constexpr int F_NAMCAP = 32;   // NAMs a read may have in the fused kernel
// hits of a read written by the lookup itself into the read's own region of
// the hit array; reads with more go to the pool and a second pass. 128 holds
// the hits of non-repetitive reads up to 300 bp (~100 randstrobes), so that
// only repetitive reads take the pool's global counters (one atomic each)
constexpr uint32_t HIT_REGION = 128;

constexpr uint32_t BUCKET_LINE = 8;  // uint4 per bucket line

// launch wrappers (kernels.hip)
void launch_bucket_lines(const uint4 *rs, const uint64_t *starts, int bits, uint4 *lines, hipStream_t st);
void launch_reset_ctl(DevCtl *ctl, hipStream_t st);
void launch_seed(const DevBatch &b, const ReadSet &set, const SeedParams &sp, int want_syncmers,
                 int long_reads, hipStream_t st);
// rescue_mode: 0 = the caller's decision (nonrepetitive fraction < 0.7 or no
// hits), 1 = never rescue, 2 = always rescue
void launch_lookup(const DevBatch &b, const ReadSet &set, const DevIndex &ix, int rescue_mode,
                   hipStream_t st);
void launch_hits(const DevBatch &b, const ReadSet &set, const DevIndex &ix, hipStream_t st);
void launch_rescue_count(const DevBatch &b, const ReadSet &set, const DevIndex &ix,
                         uint32_t rescue_cutoff, uint32_t hard_cutoff, hipStream_t st);
void launch_rescue_hits(const DevBatch &b, const ReadSet &set, const DevIndex &ix, hipStream_t st);
void launch_merge(const DevBatch &b, const ReadSet &set, int rescue, hipStream_t st);
// fused kernel over queue[0 .. count); lmax in {160, 320}, wpb in {2, 4, 8};
// returns false for an unsupported combination
bool launch_fused(const DevBatch &b, const DevIndex &ix, const SeedParams &sp, uint32_t rescue_cutoff,
                  uint32_t hard_cutoff, const uint32_t *queue, uint32_t count, int lmax, int wpb,
                  unsigned grid, uint32_t chunk, unsigned int *head, hipStream_t st);
void launch_compact(const DevBatch &b, int fused, hipStream_t st);
// the control block to pinned host memory (dst: the device's address of it)
void launch_ctl_to_host(const DevCtl *src, DevCtl *dst, hipStream_t st);
void launch_find_flat(const DevIndex &ix, const uint64_t *hashes, uint64_t n, uint64_t *positions,
                      hipStream_t st);
// exclusive scan of u32 counts into u64 offsets[n+1] (offsets[n] = total);
// tmp must hold scan_tmp_words(n) u64.
size_t scan_tmp_words(uint64_t n);
void launch_scan(const uint32_t *in, uint64_t *out, uint64_t n, uint64_t *tmp, hipStream_t st);

}  // namespace gis
