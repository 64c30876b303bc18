// gfx950 kernels of the strobealign seeding path.
//
// One wavefront (64 lanes) owns one read in every per-read kernel. Lanes
// work on positions, windows, seeds or hits of that read in parallel and the
// wave synchronises through LDS or through its own scratch regions. Reads are
// independent: variable-size per-read outputs are placed by scans of per-read
// counts, and waves only meet in the fused kernel's spill list, whose order
// never changes a read's result.
//
// Two forms share the device functions below:
//  - per-stage: k_seed -> k_lookup -> k_hits -> k_rescue_count ->
//    k_rescue_hits -> k_merge (normal, rescue), one launch per stage;
//  - fused: k_fused runs the whole path of a read in one wave out of LDS and
//    hands reads that do not fit its LDS budget (long reads, many seeds,
//    many hits) to the per-stage kernels through the spill list.
// Both end in a scan of the per-read NAM counts and k_compact.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"
#include "seed_dev.h"
#include "wave.h"

namespace gis {

// wave-uniform iteration over a ReadSet
__device__ __forceinline__ uint32_t set_count(const ReadSet &s) { return s.ids ? *s.n_dev : s.n; }
__device__ __forceinline__ uint32_t set_at(const ReadSet &s, uint32_t w) { return s.ids ? s.ids[w] : w; }

// ---------------------------------------------------------------------------
// seeding: canonical syncmers and query randstrobes of one read

struct SeedBufs {
    uint64_t *pk;       // bases 2 bits each, LSB first: base i at bits 2(i%32) of pk[i/32]
    uint8_t *st;        // per k-mer window: anchor state / ST_UNKNOWN / ST_INVALID
    uint64_t *sh;       // canonical s-mer hash per s-mer start
    uint64_t *sy_hash;  // syncmer list
    uint16_t *sy_pos;   // read positions fit 16 bits (reads are at most 65535 bp)
    uint64_t *nmask;    // bit i set when base i is not ACGT
};

// The passes of the syncmer scan after packing: canonical s-mer hashes,
// window states, tie replay and compaction, k-mer hashes. CLEAN: the read
// holds only ACGT bases (a wave-uniform fact), so no window needs the N
// mask and only the read's first window starts a segment.
template <bool CLEAN>
__device__ __forceinline__ uint32_t syncmer_passes(int L, const SeedParams &P, const SeedBufs &B, uint32_t cap,
                                                   uint64_t *out_sy_hash, uint64_t *out_sy_pos) {
    const int lane = (int)lane_id();
    const int k = P.k, s = P.s, w = P.w;

    // canonical s-mer hashes, one s-mer per lane (step2/randstrobes.cu:23-32);
    // an s-mer of at most 16 bases is a 32-bit code
    const int ns = L - s + 1;
    if (s <= 16) {
        for (int j0 = 0; j0 < ns; j0 += 64) {
            const int j = j0 + lane;
            if (j < ns) {
                uint64_t h = 0;
                if (CLEAN || !any_n(B.nmask, j, j + s - 1))
                    h = xxh64_u64((uint64_t)canonical_code32(B.pk, j, s, (uint32_t)P.smask));
                B.sh[j] = h;
            }
        }
    } else {
        for (int j0 = 0; j0 < ns; j0 += 64) {
            const int j = j0 + lane;
            if (j < ns) {
                uint64_t h = 0;
                if (CLEAN || !any_n(B.nmask, j, j + s - 1)) h = xxh64_u64(canonical_code(B.pk, j, s, P.smask));
                B.sh[j] = h;
            }
        }
    }
    wave_sync();

    // window states that need no history
    const int nk = L - k + 1;
    for (int p0 = 0; p0 < nk; p0 += 64) {
        const int p = p0 + lane;
        if (p < nk) {
            uint8_t st;
            if (!CLEAN && any_n(B.nmask, p, p + k - 1)) {
                st = ST_INVALID;
            } else {
                const bool seg_first =
                    (p == 0) || (!CLEAN && ((B.nmask[(p - 1) >> 6] >> ((p - 1) & 63)) & 1));
                // w = k - s + 1 is 5 in every read-length profile up to 375 bp
                st = w == 5 ? anchor_state_w<5>(B.sh, p, seg_first) : anchor_state(B.sh, p, w, seg_first);
            }
            B.st[p] = st;
        }
    }
    wave_sync();

    // resolve tied windows from the nearest anchor and compact syncmer
    // positions; a window is a syncmer when its minimum sits at offset t-1
    // (step2/randstrobes.cu:64-69)
    uint32_t nsync = 0;
    for (int p0 = 0; p0 < nk; p0 += 64) {
        const int p = p0 + lane;
        bool is_sync = false;
        if (p < nk) {
            uint8_t st = B.st[p];
            if (st == ST_UNKNOWN) {
                // window p-1 is valid and in the same segment; walk back to the
                // segment's nearest anchor (the segment's first window is one)
                int q = p - 1;
                while (q > 0 && B.st[q] == ST_UNKNOWN) q--;
                uint8_t r = B.st[q];
                if (r != ST_NONE && r >= (uint8_t)w) r = ST_NONE;  // cannot happen; keeps indexing in range
                for (int x = q + 1; x <= p; x++) r = slide(r, B.sh, x, w);
                st = r;
            }
            is_sync = st == (uint8_t)(P.t - 1);
        }
        const uint64_t bal = __ballot(is_sync);
        if (is_sync) {
            const uint32_t idx = nsync + mbcnt(bal);
            if (idx < cap) B.sy_pos[idx] = (uint16_t)p;
            if (out_sy_pos) out_sy_pos[idx] = (uint64_t)p;
        }
        nsync += (uint32_t)__popcll(bal);
    }
    wave_sync();

    // k-mer hashes over the compacted list (xxh64 of the canonical k-mer)
    const uint32_t nh = nsync < cap ? nsync : cap;
    for (uint32_t i0 = 0; i0 < nh; i0 += 64) {
        const uint32_t i = i0 + (uint32_t)lane;
        if (i < nh) {
            const uint64_t hk = xxh64_u64(canonical_code(B.pk, (int)B.sy_pos[i], k, P.kmask));
            B.sy_hash[i] = hk;
            if (out_sy_hash) out_sy_hash[i] = hk;
        }
    }
    wave_sync();
    return nsync;
}

// Canonical syncmers of one read (SyncmerIterator, step2/randstrobes.cu:18-79),
// the whole wave working on the same read. The first `cap` syncmers go to
// B.sy_hash/B.sy_pos (and out_sy_* when given); returns the total count.
__device__ __forceinline__ uint32_t seed_syncmers(const char *__restrict__ seq, int L,
                                                  const SeedParams &P, const SeedBufs &B,
                                                  uint32_t cap, uint64_t *out_sy_hash,
                                                  uint64_t *out_sy_pos) {
    const int lane = (int)lane_id();

    // 2-bit packed bases and the N mask from three ballots per 64 bases
    // (step2/randstrobes.cu:21-22, 70-76)
    int nch = 0;
    uint64_t nany = 0;
    for (int b0 = 0; b0 < L; b0 += 64, nch++) {
        const int i = b0 + lane;
        int c = 4;
        if (i < L) c = nt4((unsigned char)seq[i]);
        const bool ok = i < L && c < 4;
        const uint64_t lo = __ballot(ok && (c & 1));
        const uint64_t hi = __ballot(ok && (c & 2));
        const uint64_t m = __ballot(i < L && c > 3);
        nany |= m;
        if (lane == 0) {
            B.nmask[b0 >> 6] = m;
            B.pk[2 * nch] = spread32((uint32_t)lo) | (spread32((uint32_t)hi) << 1);
        } else if (lane == 1) {
            B.pk[2 * nch + 1] = spread32((uint32_t)(lo >> 32)) | (spread32((uint32_t)(hi >> 32)) << 1);
        }
    }
    if (lane == 0) B.pk[2 * nch] = 0;  // pad word read by windows ending in the last word
    wave_sync();
    // the ballots are wave-uniform, so this branch is too
    return nany == 0 ? syncmer_passes<true>(L, P, B, cap, out_sy_hash, out_sy_pos)
                     : syncmer_passes<false>(L, P, B, cap, out_sy_hash, out_sy_pos);
}

// Randstrobe number j of a read's query list in randstrobes_query order
// (step2/randstrobes.cu:195-241): j < n_fwd is the forward randstrobe of
// strobe1 = syncmer j, j >= n_fwd the reverse-complement one of strobe1 =
// syncmer n - 1 - (j - n_fwd). The two directions share one code path (lanes
// of both run together without divergence): the reverse-complement pass walks
// the syncmer list backwards, and its positions are mirrored, L - k - pos.
// Candidates lie at index offsets w_min..w_max (RandstrobeIterator::get,
// step2/randstrobes.cu:136-159); positions grow with the index, so the
// reference's break at the first candidate beyond max_dist is the same as
// skipping every candidate beyond it.
struct Rs1 {
    uint64_t h;
    uint32_t s, e;
};

__device__ __forceinline__ Rs1 randstrobe_one(const SeedBufs &B, uint32_t n, uint32_t n_fwd, uint32_t j, int L,
                                              const SeedParams &P) {
    const bool rc = j >= n_fwd;
    const uint32_t i = rc ? j - n_fwd : j;
    const uint32_t last = n - 1;
    const uint32_t k = (uint32_t)P.k;
    const uint32_t mirror = (uint32_t)L - k;  // a syncmer starts at most at L - k
    const uint32_t a = rc ? last - i : i;
    const uint32_t p1 = rc ? mirror - B.sy_pos[a] : B.sy_pos[a];
    const uint64_t h1 = B.sy_hash[a];
    // 32-bit arithmetic in the candidate loop: positions are below 2^32, so
    // the limit clamps there; the cost is at most 64; the high halves of the
    // hashes only count when q has high bits (q = 2^c - 1, c = 8 by default)
    const uint64_t maxp64 = (uint64_t)p1 + P.max_dist_u;
    const uint32_t maxp = maxp64 > 0xffffffffull ? 0xffffffffu : (uint32_t)maxp64;
    const uint32_t qlo = (uint32_t)P.q, qhi = (uint32_t)(P.q >> 32);
    const uint32_t h1lo = (uint32_t)h1, h1hi = (uint32_t)(h1 >> 32);
    const uint32_t *hw = reinterpret_cast<const uint32_t *>(B.sy_hash);
    uint32_t best = a, pbest = p1, minv = 0xffffffffu;
#pragma unroll 1
    for (uint32_t d = P.w_min; d <= P.w_max; d++) {
        if (i + d <= last) {
            const uint32_t x = rc ? a - d : a + d;
            uint32_t c = (uint32_t)__popc((h1lo ^ hw[2 * x]) & qlo);
            if (qhi) c += (uint32_t)__popc((h1hi ^ hw[2 * x + 1]) & qhi);
            const uint32_t px = rc ? mirror - B.sy_pos[x] : B.sy_pos[x];
            if (px <= maxp && c < minv) {
                minv = c;
                best = x;
                pbest = px;
            }
        }
    }
    Rs1 o;
    o.h = h1 + B.sy_hash[best];
    o.s = p1;
    o.e = pbest + k;
    return o;
}

// randstrobes_query order: forward (strobe1 ascending) then reverse complement,
// one randstrobe per lane
__device__ __forceinline__ uint32_t seed_one(const char *__restrict__ seq, int L, const SeedParams &P,
                                             const SeedBufs &B, uint64_t *out_sy_hash,
                                             uint64_t *out_sy_pos, uint64_t *rs_hash,
                                             uint32_t *rs_span,
                                             uint32_t *n_sync_out) {
    const uint32_t n = seed_syncmers(seq, L, P, B, 0xffffffffu, out_sy_hash, out_sy_pos);
    const uint32_t n_fwd = ((uint32_t)L >= P.w_max && n > P.w_min) ? n - P.w_min : 0;
    const uint32_t nrs = 2 * n_fwd;
    for (uint32_t j0 = 0; j0 < nrs; j0 += 64) {
        const uint32_t j = j0 + lane_id();
        if (j < nrs) {
            const Rs1 o = randstrobe_one(B, n, n_fwd, j, L, P);
            rs_hash[j] = o.h;
            rs_span[j] = span_pack(o.s, o.e);
        }
    }
    *n_sync_out = n;
    return nrs;
}

struct alignas(8) SeedLds {
    uint64_t sh[LMAX_LDS];
    uint64_t sy_hash[LMAX_LDS];
    uint64_t nmask[LMAX_LDS / 64 + 2];
    uint64_t pk[2 * (LMAX_LDS / 64) + 1];
    uint16_t sy_pos[LMAX_LDS];
    uint8_t st[LMAX_LDS];
};

__host__ __device__ inline uint64_t align8(uint64_t x) { return (x + 7) & ~7ULL; }

// global scratch layout of a long read (long_bytes() on the host matches it)
__device__ __forceinline__ SeedBufs long_bufs(uint8_t *base, uint64_t L) {
    SeedBufs B;
    B.sh = (uint64_t *)base;
    B.sy_hash = (uint64_t *)(base + 8 * L);
    B.nmask = (uint64_t *)(base + 16 * L);
    B.sy_pos = (uint16_t *)(base + 16 * L + 8 * (L / 64 + 2));
    // the 8 (2 ceil(L / 64) + 1) bytes of packed bases fit the L bytes given to them (L > 320)
    B.pk = (uint64_t *)(base + 16 * L + 8 * (L / 64 + 2) + align8(4 * L));
    B.st = (uint8_t *)B.pk + align8(L);
    return B;
}

// Seeding parameters of the read-length profiles up to 250 bp (k = 20,
// s = 16, t = 3): kernels instantiated with K20 take them as compile-time
// constants, which shortens the seeding code and frees the scalar registers
// that would hold them. The host launches that instantiation when the
// index's parameters match (is_k20_profile).
__device__ __forceinline__ void pin_k20_profile(SeedParams &P) {
    P.k = 20;
    P.s = 16;
    P.t = 3;
    P.w = 5;
    P.kmask = (1ull << 40) - 1;
    P.smask = 0xffffffffull;
}

static inline bool is_k20_profile(const SeedParams &sp) { return sp.k == 20 && sp.s == 16 && sp.t == 3; }

template <bool LONG, bool K20>
__global__ __launch_bounds__(64 * SEED_WPB) __attribute__((amdgpu_waves_per_eu(6))) void k_seed(DevBatch b, ReadSet set, SeedParams P,
                                                        int want_sy) {
    if (K20) pin_k20_profile(P);
    __shared__ SeedLds lds[LONG ? 1 : SEED_WPB];
    const int wid = (int)uni32(threadIdx.x >> 6);
    const uint32_t cnt = set_count(set);
    for (uint32_t w = wave_index(SEED_WPB); w < cnt; w += gridDim.x * SEED_WPB) {
        const uint32_t r = set.ids ? sload(set.ids + w) : w;
        const int L = (int)sload(b.read_len + r);
        if ((L > LMAX_LDS) != LONG) continue;
        SeedBufs B;
        if (LONG) {
            B = long_bufs(b.long_scratch + sload(b.long_off + r), (uint64_t)L);
        } else {
            SeedLds &S = lds[wid];
            B = SeedBufs{S.pk, S.st, S.sh, S.sy_hash, S.sy_pos, S.nmask};
        }
        const uint64_t so = want_sy ? sload(b.sy_off + r) : 0;
        uint64_t *osh = want_sy ? b.sy_hash + so : nullptr;
        uint64_t *osp = want_sy ? b.sy_pos + so : nullptr;
        const uint64_t ro = sload(b.rs_off + r);
        uint32_t nsync = 0;
        const uint32_t nrs = seed_one(b.seqs + sload(b.read_off + r), L, P, B, osh, osp, b.rs_hash + ro,
                                      b.rs_span + ro, &nsync);
        if (lane_id() == 0) {
            b.rs_cnt[r] = nrs;
            b.rs_nfwd[r] = nrs / 2;
            if (want_sy) b.sy_cnt[r] = nsync;
        }
        wave_sync_mem();
    }
}

// ---------------------------------------------------------------------------
// index access (StrobemerIndex, step2/index.hpp:57-147)

__device__ __forceinline__ uint64_t rec_hash(const DevIndex &ix, uint64_t x) {
    return reinterpret_cast<const uint64_t *>(ix.rs)[2 * x];
}

// first index of the bucket range holding key, or GIS_NOT_FOUND
// (StrobemerIndex::find, step2/index.hpp:57-81: linear scan of a small bucket,
// lower_bound otherwise; both return the first equal entry). *scan is the
// record count of the byte model: the records a linear scan reads up to the
// first hash >= key (all of them when there is none), whichever search ran.
__device__ __forceinline__ uint64_t dev_find_in(const DevIndex &ix, uint64_t key, uint64_t s, uint64_t e,
                                               uint32_t *scan) {
    uint32_t pr = 0;
    uint64_t res = GIS_NOT_FOUND;
    if (s == e) {
    } else if (e - s < 8) {
        for (uint64_t x = s; x < e; x++) {
            const uint64_t h = rec_hash(ix, x);
            pr++;
            if (h >= key) {
                res = h == key ? x : GIS_NOT_FOUND;
                break;
            }
        }
    } else {
        uint64_t lo = s, cnt = e - s;
        while (cnt > 0) {
            const uint64_t step = cnt >> 1, it = lo + step;
            pr++;
            if (rec_hash(ix, it) < key) { lo = it + 1; cnt -= step + 1; }
            else cnt = step;
        }
        if (lo < e && rec_hash(ix, lo) == key) res = lo;
        pr = (uint32_t)(lo - s) + (lo < e ? 1u : 0u);
    }
    *scan = pr;
    return res;
}

__device__ __forceinline__ uint64_t dev_find(const DevIndex &ix, uint64_t key, uint32_t *scan) {
    const uint32_t bk = (uint32_t)(key >> (64 - ix.bits));
    return dev_find_in(ix, key, ix.starts[bk], ix.starts[bk + 1], scan);
}

__device__ __forceinline__ uint64_t dev_find(const DevIndex &ix, uint64_t key) {
    uint32_t pr;
    return dev_find(ix, key, &pr);
}

// is_filtered: hash[pos] == hash[pos + filter_cutoff], past-the-end reads as ~0
// (step2/index.hpp:83-93)
__device__ __forceinline__ bool dev_filtered(const DevIndex &ix, uint64_t pos, uint64_t key) {
    const uint64_t x = pos + ix.filter_cutoff;
    const uint64_t hx = x < ix.n ? rec_hash(ix, x) : UINT64_MAX;
    return hx == key;
}

// get_count: length of the equal-hash run at pos inside its bucket
// (step2/index.hpp:115-147)
__device__ __forceinline__ uint32_t dev_get_count(const DevIndex &ix, uint64_t pos) {
    const uint64_t key = rec_hash(ix, pos);
    const uint32_t bk = (uint32_t)(key >> (64 - ix.bits));
    const uint64_t pe = ix.starts[bk + 1];
    if (pe - pos < 8) {
        uint32_t c = 1;
        for (uint64_t x = pos + 1; x < pe; x++) {
            if (rec_hash(ix, x) == key) c++;
            else break;
        }
        return c;
    }
    uint64_t lo = pos, cnt = pe - pos;
    while (cnt > 0) {
        const uint64_t step = cnt >> 1, it = lo + step;
        if (!(key < rec_hash(ix, it))) { lo = it + 1; cnt -= step + 1; }
        else cnt = step;
    }
    return (uint32_t)(lo - pos);
}

// Walk the equal-hash run at pos (add_to_hits_per_ref, step2/nam.cu:45-62):
// keep entry e iff |query span - ref span| <= the running minimum. Counts the
// kept hits, and writes them when out != nullptr; *run receives the run's
// length.
__device__ __forceinline__ uint32_t walk_run(const DevIndex &ix, uint64_t pos, uint64_t key,
                                             int qs, int qe, uint32_t strand_bit, DevHit *out,
                                             uint32_t *run = nullptr) {
    const int qspan = qe - qs;
    int min_diff = 0x7fffffff;
    uint32_t c = 0, rec = 0;
    for (uint64_t x = pos; x < ix.n; x++) {
        const uint4 v = ix.rs[x];
        const uint64_t h = (uint64_t)v.x | ((uint64_t)v.y << 32);
        if (h != key) break;
        rec++;
        const int rspan = (int)(v.w & 0xffu) + ix.k;
        int d = qspan - rspan;
        d = d < 0 ? -d : d;
        if (d <= min_diff) {
            if (out) {
                DevHit hh;
                hh.ref_strand = strand_bit | (v.w >> 8);
                hh.ref_start = (int32_t)v.z;
                hh.ref_end = (int32_t)v.z + rspan;
                hh.query_start = (uint16_t)qs;
                hh.query_end = (uint16_t)qe;
                out[c] = hh;
            }
            c++;
            min_diff = d;
        }
    }
    if (run) *run = rec;
    return c;
}

// The same walk by the whole wave, 64 records per step, for the long runs of
// the rescue path (up to hard_cutoff records): the running minimum of
// add_to_hits_per_ref is the minimum of all earlier differences (an entry
// that is not kept does not lower it), so entry e is kept iff its difference
// is at most the exclusive prefix minimum at e. Returns the kept count (the
// same on every lane) and writes the kept hits in run order when out !=
// nullptr. Every lane calls it with the same arguments.
constexpr uint32_t WALK_WAVE_MIN = 32;  // runs longer than this take the wave walk

__device__ __forceinline__ uint32_t walk_run_wave(const DevIndex &ix, uint64_t pos, uint64_t key, int qs, int qe,
                                                  uint32_t strand_bit, DevHit *out) {
    const uint32_t lane = lane_id();
    const int qspan = qe - qs;
    int carry = 0x7fffffff;
    uint32_t kept = 0;
    for (uint64_t x0 = pos; x0 < ix.n; x0 += 64) {
        const uint64_t x = x0 + lane;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        bool in = false;
        if (x < ix.n) {
            v = ix.rs[x];
            in = ((uint64_t)v.x | ((uint64_t)v.y << 32)) == key;
        }
        // equal hashes are contiguous: the run ends at the first other record
        const uint64_t other = __ballot(!in);
        const uint32_t len = other ? (uint32_t)__builtin_ctzll(other) : 64u;
        const bool valid = lane < len;
        const int rspan = (int)(v.w & 0xffu) + ix.k;
        int d = qspan - rspan;
        d = d < 0 ? -d : d;
        int incl = valid ? d : 0x7fffffff;
        for (int sh = 1; sh < 64; sh <<= 1) {
            const int y = __shfl_up(incl, sh, 64);
            if (lane >= (uint32_t)sh) incl = y < incl ? y : incl;
        }
        int excl = __shfl_up(incl, 1, 64);
        if (lane == 0) excl = 0x7fffffff;
        excl = carry < excl ? carry : excl;
        const bool keep = valid && d <= excl;
        const uint64_t km = __ballot(keep);
        if (out && keep) {
            DevHit hh;
            hh.ref_strand = strand_bit | (v.w >> 8);
            hh.ref_start = (int32_t)v.z;
            hh.ref_end = (int32_t)v.z + rspan;
            hh.query_start = (uint16_t)qs;
            hh.query_end = (uint16_t)qe;
            out[kept + mbcnt(km)] = hh;
        }
        kept += (uint32_t)__popcll(km);
        const int last = (int)readlane_u32((uint32_t)incl, 63);
        carry = last < carry ? last : carry;
        if (len < 64) break;
    }
    return kept;
}

// ---------------------------------------------------------------------------
// Probe of one query randstrobe for the per-stage lookup. The result is the
// reference's (StrobemerIndex::find + is_filtered + the run walk of
// add_to_hits_per_ref), reached with fewer dependent memory steps: the
// bucket's line (DevIndex::lines) holds its bounds and first records, so one
// line fetch serves a bucket of at most PROBE_SMALL records (nearly all of
// them: buckets hold about 2.4 records on average). The first equal record,
// the run length (equal hashes never cross a bucket), the filter test and the
// kept hits then follow from registers. Larger buckets take the reference's
// search and the separate filter probe in the table.

constexpr uint32_t PROBE_SMALL = BUCKET_LINE - 1;  // 7: 99.6% of buckets at 2.4 records per bucket

// a table record as the probe reads it (gis_ref_randstrobe)
struct alignas(16) IdxRec {
    uint64_t hash;
    uint32_t position;
    uint32_t packed;  // ref_index << 8 | strobe2 offset
};

struct Probe {
    uint64_t key, s, e, pos;
    IdxRec r[PROBE_SMALL];  // records s, s + 1, ... of a small bucket
    uint32_t scan;          // records a linear scan reads to the first hash >= key (byte model)
    uint32_t run;           // length of the equal-hash run of a kept seed (byte model)
    uint32_t hc;            // kept hits (0 when absent or filtered)
    uint32_t keep;          // small bucket: bit j = record s + j is a kept hit
    bool found, filtered;
};

__device__ __forceinline__ bool probe_small(const Probe &p) { return p.e - p.s <= PROBE_SMALL; }

// Bucket bounds (StrobemerIndex::find, step2/index.hpp:57-62) and the
// bucket's first PROBE_SMALL records, all from the bucket's line: one cache
// line, independent loads, no dependent second fetch. Records past the
// bucket's size are zeros and never compared. A size of 2^32 - 1 or more is
// clamped in the line; such buckets are large, and large buckets search the
// table between the starts[] bounds.
__device__ __forceinline__ void probe_line(const DevIndex &ix, Probe &p) {
    const uint32_t bk = (uint32_t)(p.key >> (64 - ix.bits));
    const uint4 *line = ix.lines + (uint64_t)bk * BUCKET_LINE;
    const uint4 h = line[0];
    const IdxRec *recs = reinterpret_cast<const IdxRec *>(line + 1);
#pragma unroll
    for (uint32_t j = 0; j < PROBE_SMALL; j++) p.r[j] = recs[j];
    p.s = (uint64_t)h.x | ((uint64_t)h.y << 32);
    p.e = p.s + h.z;
}

// Search, filter and run walk. Small buckets: the reference's linear scan
// stops at the first record >= key, which is record s + (records below key);
// the run is the records equal to key; hash[pos + filter_cutoff] == key
// exactly when the run is longer than filter_cutoff (past the bucket the
// hashes differ, past the table the reference reads ~0).
// the first equal record of a small bucket from the line's records:
// records below the key, records equal to it
__device__ __forceinline__ void probe_small_find(Probe &p, uint32_t &lt, uint32_t &eq) {
    const uint32_t sz = (uint32_t)(p.e - p.s);
    lt = 0;
    eq = 0;
#pragma unroll
    for (uint32_t j = 0; j < PROBE_SMALL; j++) {
        const uint64_t h = p.r[j].hash;
        if (j < sz) {
            lt += h < p.key ? 1u : 0u;
            eq += h == p.key ? 1u : 0u;
        }
    }
    p.scan = lt < sz ? lt + 1 : sz;
    p.found = eq > 0;
    p.pos = p.found ? p.s + lt : GIS_NOT_FOUND;
}

__device__ __forceinline__ void probe_finish(const DevIndex &ix, Probe &p, int qs, int qe) {
    p.hc = 0;
    p.keep = 0;
    p.run = 0;
    p.filtered = false;
    if (probe_small(p)) {
        uint32_t lt, eq;
        probe_small_find(p, lt, eq);
        if (!p.found) return;
        p.filtered = eq > ix.filter_cutoff ||
                     (p.key == UINT64_MAX && p.pos + ix.filter_cutoff >= ix.n);
        if (p.filtered) return;
        p.run = eq;
        const int qspan = qe - qs;
        int min_diff = 0x7fffffff;
#pragma unroll
        for (uint32_t j = 0; j < PROBE_SMALL; j++) {
            if (j >= lt && j < lt + eq) {
                int d = qspan - ((int)(p.r[j].packed & 0xffu) + ix.k);
                d = d < 0 ? -d : d;
                if (d <= min_diff) {
                    p.keep |= 1u << j;
                    p.hc++;
                    min_diff = d;
                }
            }
        }
        return;
    }
    p.pos = p.e - p.s < 0xffffffffull ? dev_find_in(ix, p.key, p.s, p.e, &p.scan) : dev_find(ix, p.key, &p.scan);
    p.found = p.pos != GIS_NOT_FOUND;
    if (!p.found) return;
    p.filtered = dev_filtered(ix, p.pos, p.key);
    if (p.filtered) return;
    p.hc = walk_run(ix, p.pos, p.key, qs, qe, 0u, nullptr, &p.run);
}

// the kept hits of a probe, in run order
__device__ __forceinline__ void probe_emit(const DevIndex &ix, const Probe &p, int qs, int qe,
                                           uint32_t strand_bit, DevHit *out) {
    if (!probe_small(p)) {
        walk_run(ix, p.pos, p.key, qs, qe, strand_bit, out);
        return;
    }
    uint32_t c = 0;
#pragma unroll
    for (uint32_t j = 0; j < PROBE_SMALL; j++) {
        if ((p.keep >> j) & 1u) {
            DevHit hh;
            hh.ref_strand = strand_bit | (p.r[j].packed >> 8);
            hh.ref_start = (int32_t)p.r[j].position;
            hh.ref_end = (int32_t)p.r[j].position + (int)(p.r[j].packed & 0xffu) + ix.k;
            hh.query_start = (uint16_t)qs;
            hh.query_end = (uint16_t)qe;
            out[c++] = hh;
        }
    }
}

// ---------------------------------------------------------------------------
// per-stage: lookup + filter + hit count per randstrobe; nonrepetitive
// fraction and the rescue decision per read (find_nams, step2/nam.cu:553-732,
// and the caller's test, step2/main.cu:3209-3216). One randstrobe per lane,
// 64 per pass; a 150 bp read's ~42 randstrobes fill one pass. While the
// read's hits fit its region of the hit array they are written there in seed
// order; a read with more hits takes a place in the pool, and k_hits writes
// them.
//
// Write traffic is part of this kernel's bound, so a one-pass read (at most
// 64 randstrobes) keeps its per-randstrobe results (position, kept-hit count)
// in registers and stores them only when a later kernel reads them: when the
// read is rescued (k_rescue_count, k_rescue_hits) or its hits go to the pool
// (k_hits). Longer reads, and the fused kernel's spilled reads, store them
// in every pass. Wave sums and scans use DPP (52 VGPRs).
__global__ __launch_bounds__(64 * WPB) __attribute__((amdgpu_waves_per_eu(6))) void k_lookup(
    DevBatch b, ReadSet set, DevIndex ix, int rescue_mode, int keep_spill) {
    const uint32_t lane = lane_id();
    const uint32_t cnt = set_count(set);
    for (uint32_t w = wave_index(WPB); w < cnt; w += gridDim.x * WPB) {
        const uint32_t r = set.ids ? sload(set.ids + w) : w;
        const uint32_t n = sload(b.rs_cnt + r);
        const uint64_t base = sload(b.rs_off + r);
        const uint32_t half = sload(b.rs_nfwd + r);
        // results stay in registers for a one-pass read the decision below
        // does not hand to a later kernel
        const bool defer = n <= 64 && !keep_spill && rescue_mode != 2;
        uint32_t found = 0, good = 0, hits = 0, bytes = 0, streams = 0;
        uint64_t pos0 = GIS_NOT_FOUND;
        uint32_t hc0 = 0;
        DevHit *region = b.hits + (uint64_t)r * HIT_REGION;
        // the regions must lie inside the hit array (else every read pools)
        const bool region_ok = (uint64_t)(r + 1) * HIT_REGION <= b.hit_cap;
        for (uint32_t i0 = 0; i0 < n; i0 += 64) {
            // straight-line loads (lanes past the end repeat the last item)
            const uint32_t i = i0 + lane;
            const bool act = i < n;
            const uint64_t x = base + (act ? i : n - 1);
            Probe P;
            P.key = b.rs_hash[x];
            const uint32_t span = b.rs_span[x];
            const int qs = (int)span_start(span), qe = (int)span_end(span);
            probe_line(ix, P);
            // the scheduler may not hoist uses of the line between its loads
            __builtin_amdgcn_sched_barrier(0);
            uint32_t by = 0, sb = 0;
            if (act) {
                probe_finish(ix, P, qs, qe);
                // the table bytes of the search: the bucket bounds (16), the
                // records a linear scan reads to the first hash >= key; a
                // found seed's filter probe (8), and a kept one's run past
                // its first record plus the hash that ends it: 16 per run
                // record in all. Beside them the streams: the query in (12)
                // and, where stored, the result out (12).
                by = 16 + 16 * P.scan;
                if (P.found) by += P.filtered ? 8u : 16u * P.run;
                sb = 12;
                if (defer) {
                    pos0 = P.pos;
                    hc0 = P.hc;
                } else {
                    b.pos[base + i] = P.pos;
                    b.hcnt[base + i] = P.hc;
                    sb += 12;
                }
            } else {
                P.hc = 0;
                P.found = false;
                P.filtered = false;
            }
            found += wave_count(act && P.found);
            good += wave_count(act && P.found && !P.filtered);
            const uint32_t incl = wave_incl_scan_dpp(P.hc);
            const uint32_t chunk_hits = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
            // while the read's hits fit its region, write them there in seed
            // order: 16 per hit
            if (region_ok && hits + chunk_hits <= HIT_REGION) {
                if (P.hc > 0)
                    probe_emit(ix, P, qs, qe, i >= half ? 0x80000000u : 0u, region + hits + (incl - P.hc));
                sb += 16 * P.hc;
            }
            hits += chunk_hits;
            bytes += wave_sum_dpp(by);
            streams += wave_sum_dpp(sb);
        }
        const float frac = found > 0 ? (float)good / (float)found : 1.0f;
        const bool rescue = rescue_mode == 2 || (rescue_mode == 0 && ((double)frac < 0.7 || hits == 0));
        const bool pooled = !rescue && hits > 0 && (hits > HIT_REGION || !region_ok);
        if (defer && (rescue || pooled)) {
            if (lane < n) {
                b.pos[base + lane] = pos0;
                b.hcnt[base + lane] = hc0;
            }
            streams += 12 * n;
        }
        if (lane == 0) {
            b.lk_bytes[r] = bytes;
            b.lk_stream[r] = streams;
            b.frac[r] = frac;
            const uint8_t keep = keep_spill ? (uint8_t)(b.flags[r] & FLAG_SPILL) : (uint8_t)0;
            b.flags[r] = (uint8_t)(keep | (rescue ? GIS_FLAG_RESCUED : 0) |
                                   (n == 0 ? GIS_FLAG_NO_RANDSTROBES : 0));
            b.hit_total[r] = rescue ? 0u : hits;  // rescued reads discard these hits
            b.rhit_total[r] = 0;
            // hits beyond the read's region: a place in the pool, and k_hits
            // writes them
            uint64_t off = (uint64_t)r * HIT_REGION;
            if (pooled) {
                off = (uint64_t)b.n_reads * HIT_REGION + atomicAdd(&b.ctl->hit_pool, (unsigned long long)hits);
                b.sel_ids[2ull * b.n_reads + atomicAdd(&b.ctl->n_sel[2], 1u)] = r;
            }
            b.hit_off[r] = off;
        }
    }
}

// end of the hit pool, for the capacity checks of the kernels that follow
__global__ void k_pool_total(DevBatch b) {
    if (threadIdx.x == 0) {
        const uint64_t end = (uint64_t)b.n_reads * HIT_REGION + b.ctl->hit_pool;
        b.hit_off[b.n_reads] = end;
        b.ctl->tot[0] = end;
    }
}

// write the hits of the reads whose hits did not fit their region into the
// pool (same walk as k_lookup); runs over the pooled reads' list
// capacity check of a hit array against the scan total (wave-uniform)
__device__ __forceinline__ bool over_capacity(const uint64_t *off, uint32_t n, uint64_t cap,
                                              DevCtl *ctl, unsigned bit) {
    if (off[n] <= cap) return false;
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&ctl->overflow, bit);
    return true;
}

__global__ __launch_bounds__(64 * WPB) void k_hits(DevBatch b, ReadSet set, DevIndex ix) {
    const uint32_t lane = lane_id();
    if (over_capacity(b.hit_off, b.n_reads, b.hit_cap, b.ctl, OVF_HITS)) return;
    const uint32_t cnt = set_count(set);
    for (uint32_t w = wave_index(WPB); w < cnt; w += gridDim.x * WPB) {
        const uint32_t r = set.ids ? sload(set.ids + w) : w;
        if (sload(b.hit_total + r) == 0) continue;
        const uint32_t n = sload(b.rs_cnt + r);
        const uint64_t base = sload(b.rs_off + r);
        const uint32_t half = sload(b.rs_nfwd + r);
        DevHit *out = b.hits + sload(b.hit_off + r);
        uint32_t carry = 0;
        for (uint32_t i0 = 0; i0 < n; i0 += 64) {
            const uint32_t i = i0 + lane;
            const uint32_t hc = i < n ? b.hcnt[base + i] : 0u;
            const uint32_t ex = wave_excl_scan(hc);
            if (hc > 0)
                walk_run(ix, b.pos[base + i], b.rs_hash[base + i], (int)span_start(b.rs_span[base + i]),
                         (int)span_end(b.rs_span[base + i]), i >= half ? 0x80000000u : 0u, out + carry + ex);
            carry += wave_sum(hc);
        }
    }
}

// ---------------------------------------------------------------------------
// rescue (find_nams_rescue, step2/nam.cu:927-984)
//
// A seed is taken when count <= hard_cutoff and (count <= rescue_cutoff or
// fewer than 5 seeds of its strand precede it in (count, query_start,
// query_end) order): exactly the prefix the reference's sorted loop keeps
// (step2/nam.cu:957-972). Taken seeds walk their runs in query_start order,
// which is the seeds' own order within a strand.

__device__ __forceinline__ bool rh_less(uint32_t ca, uint32_t qa, uint32_t ea, uint32_t cb,
                                        uint32_t qb, uint32_t eb) {
    return ca < cb || (ca == cb && (qa < qb || (qa == qb && ea < eb)));
}

__global__ __launch_bounds__(64 * WPB) void k_rescue_count(DevBatch b, ReadSet set, DevIndex ix,
                                                           uint32_t rescue_cutoff,
                                                           uint32_t hard_cutoff) {
    const uint32_t lane = lane_id();
    const uint32_t cnt = set_count(set);
    for (uint32_t w = wave_index(WPB); w < cnt; w += gridDim.x * WPB) {
        const uint32_t r = set.ids ? sload(set.ids + w) : w;
        if (!(sload(b.flags + r) & GIS_FLAG_RESCUED)) continue;
        const uint32_t n = sload(b.rs_cnt + r);
        const uint64_t base = sload(b.rs_off + r);
        const uint32_t half = sload(b.rs_nfwd + r);
        for (uint32_t i0 = 0; i0 < n; i0 += 64) {
            const uint32_t i = i0 + lane;
            if (i < n) {
                const uint64_t pos = b.pos[base + i];
                b.hcnt[base + i] = pos != GIS_NOT_FOUND ? dev_get_count(ix, pos) : 0u;
            }
        }
        wave_sync_mem();
        uint32_t total = 0;
        for (uint32_t i0 = 0; i0 < n; i0 += 64) {
            const uint32_t i = i0 + lane;
            uint32_t rh = 0;
            bool coop = false;
            if (i < n) {
                const uint32_t c = b.hcnt[base + i];
                bool taken = false;
                if (c > 0 && c <= hard_cutoff) {
                    taken = c <= rescue_cutoff;
                    if (!taken) {
                        const uint32_t lo = i < half ? 0 : half, hi = i < half ? half : n;
                        const uint32_t sp = b.rs_span[base + i], qs = span_start(sp), qe = span_end(sp);
                        uint32_t rank = 0;
                        for (uint32_t j = lo; j < hi && rank < 5; j++) {
                            const uint32_t cj = b.hcnt[base + j];
                            if (cj == 0) continue;
                            const uint32_t spj = b.rs_span[base + j];
                            rank += rh_less(cj, span_start(spj), span_end(spj), c, qs, qe) ? 1u : 0u;
                        }
                        taken = rank < 5;
                    }
                }
                if (taken && c <= WALK_WAVE_MIN)
                    rh = walk_run(ix, b.pos[base + i], b.rs_hash[base + i], (int)span_start(b.rs_span[base + i]),
                                  (int)span_end(b.rs_span[base + i]), 0u, nullptr);
                coop = taken && c > WALK_WAVE_MIN;
            }
            // long runs: the whole wave walks each one
            uint64_t cm = __ballot(coop);
            while (cm) {
                const uint32_t l = (uint32_t)__builtin_ctzll(cm);
                cm &= cm - 1;
                const uint64_t xl = base + i0 + l;
                const uint32_t k = walk_run_wave(ix, b.pos[xl], b.rs_hash[xl], (int)span_start(b.rs_span[xl]),
                                                 (int)span_end(b.rs_span[xl]), 0u, nullptr);
                if (lane == l) rh = k;
            }
            if (i < n) b.rcnt[base + i] = rh;
            total += wave_sum(rh);
        }
        if (lane == 0) b.rhit_total[r] = total;
        wave_sync_mem();
    }
}

__global__ __launch_bounds__(64 * WPB) void k_rescue_hits(DevBatch b, ReadSet set, DevIndex ix) {
    const uint32_t lane = lane_id();
    if (blockIdx.x == 0 && threadIdx.x == 0) b.ctl->tot[1] = b.rhit_off[b.n_reads];
    if (over_capacity(b.rhit_off, b.n_reads, b.rhit_cap, b.ctl, OVF_RHITS)) return;
    const uint32_t cnt = set_count(set);
    for (uint32_t w = wave_index(WPB); w < cnt; w += gridDim.x * WPB) {
        const uint32_t r = set.ids ? sload(set.ids + w) : w;
        if (sload(b.rhit_total + r) == 0) continue;
        const uint32_t n = sload(b.rs_cnt + r);
        const uint64_t base = sload(b.rs_off + r);
        const uint32_t half = sload(b.rs_nfwd + r);
        DevHit *out = b.rhits + sload(b.rhit_off + r);
        uint32_t carry = 0;
        for (uint32_t i0 = 0; i0 < n; i0 += 64) {
            const uint32_t i = i0 + lane;
            const uint32_t rh = i < n ? b.rcnt[base + i] : 0u;
            // runs longer than WALK_WAVE_MIN (k_rescue_count left their
            // lengths in hcnt) take the wave walk, as they did there
            const bool coop = rh > 0 && b.hcnt[base + i] > WALK_WAVE_MIN;
            const uint32_t ex = wave_excl_scan(rh);
            if (rh > 0 && !coop)
                walk_run(ix, b.pos[base + i], b.rs_hash[base + i], (int)span_start(b.rs_span[base + i]),
                         (int)span_end(b.rs_span[base + i]), i >= half ? 0x80000000u : 0u, out + carry + ex);
            uint64_t cm = __ballot(coop);
            while (cm) {
                const uint32_t l = (uint32_t)__builtin_ctzll(cm);
                cm &= cm - 1;
                const uint32_t il = i0 + l;
                const uint64_t xl = base + il;
                walk_run_wave(ix, b.pos[xl], b.rs_hash[xl], (int)span_start(b.rs_span[xl]),
                              (int)span_end(b.rs_span[xl]), il >= half ? 0x80000000u : 0u,
                              out + carry + readlane_u32(ex, l));
            }
            carry += wave_sum(rh);
        }
    }
}

// ---------------------------------------------------------------------------
// NAM merge (merge_hits_into_nams, step2/nam.cu:347-513). Hits of a read
// arrive strand by strand, seeds in query_start order and each run in table
// order; within one (strand, ref) group that is (query_start, ref_start)
// order, the order the reference sorts each group into (step2/nam.cu:373).
// The hits are grouped by (strand, ref_id) with a stable sort and each group
// is merged in that order: every hit extends or is contained in the first
// open NAM (in creation order) that accepts it, or opens a new NAM
// (step2/nam.cu:391-447). A NAM whose query_end lies before the current
// query_start can never accept a hit again; when it leaves the open list does
// not change any field. Creation order inside a group is canonical order, so
// NAM c of a group is written at (NAMs of earlier groups) + c with nam_id =
// that rank.
//
// The open NAMs of the current group live one per lane; the hit is broadcast
// and a ballot picks the first accepting lane. When 64 lanes are open and
// none can be retired, the group continues on a NAM list in global scratch
// scanned 64 entries at a time (the "wide" form), retiring dead NAMs as the
// list grows.

__device__ __forceinline__ void bitonic_regs(uint64_t &key) {
    const uint32_t lane = lane_id();
    for (uint32_t size = 2; size <= 64; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            const uint64_t other = __shfl_xor(key, (int)stride, 64);
            const bool asc = (lane & size) == 0;
            const bool lower = (lane & stride) == 0;
            const uint64_t mn = key < other ? key : other, mx = key < other ? other : key;
            key = (lower == asc) ? mn : mx;
        }
    }
}

__device__ __forceinline__ void bitonic_mem(uint64_t *keys, uint32_t P) {
    const uint32_t lane = lane_id();
    for (uint32_t size = 2; size <= P; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t t0 = 0; t0 < P / 2; t0 += 64) {
                const uint32_t t = t0 + lane;
                if (t < P / 2) {
                    const uint32_t i = (t / stride) * 2 * stride + (t % stride);
                    const uint32_t j = i + stride;
                    const bool asc = (i & size) == 0;
                    const uint64_t a = keys[i], c = keys[j];
                    if ((a > c) == asc) { keys[i] = c; keys[j] = a; }
                }
            }
            wave_sync_mem();
        }
    }
}

// Stable wave radix sort of n keys (group << 32 | hit index) by the group
// field, 4 bits a pass over key_bits bits; src and alt hold n entries each.
// Returns the buffer that holds the sorted keys.
__device__ __forceinline__ uint64_t *radix_sort_groups(uint64_t *src, uint64_t *alt, uint32_t n,
                                                       uint32_t key_bits) {
    const uint32_t lane = lane_id();
    for (uint32_t shift = 32; shift < 32 + key_bits; shift += 4) {
        uint32_t off[16];
#pragma unroll
        for (int v = 0; v < 16; v++) off[v] = 0;
        for (uint32_t i0 = 0; i0 < n; i0 += 64) {
            const uint32_t i = i0 + lane;
            const uint32_t d = i < n ? (uint32_t)(src[i] >> shift) & 15u : 16u;
#pragma unroll
            for (int v = 0; v < 16; v++) off[v] += (uint32_t)__popcll(__ballot(d == (uint32_t)v));
        }
        uint32_t run = 0;
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const uint32_t c = off[v];
            off[v] = run;
            run += c;
        }
        for (uint32_t i0 = 0; i0 < n; i0 += 64) {
            const uint32_t i = i0 + lane;
            uint64_t key = 0;
            uint32_t d = 16u;
            if (i < n) {
                key = src[i];
                d = (uint32_t)(key >> shift) & 15u;
            }
            uint32_t pos = 0;
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const uint64_t bal = __ballot(d == (uint32_t)v);
                if (d == (uint32_t)v) pos = off[v] + mbcnt(bal);
                off[v] += (uint32_t)__popcll(bal);
            }
            if (i < n) alt[pos] = key;
        }
        wave_sync_mem();
        uint64_t *t = src;
        src = alt;
        alt = t;
    }
    return src;
}

constexpr uint32_t MERGE_OVERFLOW = 0xffffffffu;

struct LaneNam {
    int qs, qe, qp, rs, re, rp, n;
    uint32_t cid;
};

__device__ __forceinline__ void lane_nam_store(gis_nam *out, uint32_t idx, uint32_t rsd,
                                               const LaneNam &o) {
    gis_nam x;
    x.nam_id = (int32_t)idx;
    x.query_start = o.qs;
    x.query_end = o.qe;
    x.query_prev_hit_startpos = o.qp;
    x.ref_start = o.rs;
    x.ref_end = o.re;
    x.ref_prev_hit_startpos = o.rp;
    x.n_hits = o.n;
    x.ref_id = (int32_t)(rsd & 0x7fffffffu);
    x.score = nam_score(o.qs, o.qe, o.rs, o.re, o.n);
    x.is_rc = (uint8_t)(rsd >> 31);
    x.pad_[0] = x.pad_[1] = x.pad_[2] = 0;
    out[idx] = x;
}

__device__ __forceinline__ LaneNam from_wide(const WideNam &w) {
    return LaneNam{w.qs, w.qe, w.qp, w.rs, w.re, w.rp, (int)w.n, w.cid};
}

__device__ __forceinline__ int permute_push(int v, uint32_t dst) {
    return __builtin_amdgcn_ds_permute((int)(dst << 2), v);
}

// does open NAM o accept hit (qs, qe, rs, re), and does the hit extend it
__device__ __forceinline__ bool nam_accepts(int o_qp, int o_qe, int o_rp, int o_re, int qs, int qe,
                                            int rs, int re, bool &ext) {
    // bitwise on purpose: no short-circuit branches in the merge loops
    const bool near = (o_qp < qs) & (qs <= o_qe) & (o_rp < rs) & (rs <= o_re);
    ext = (qe > o_qe) & (re > o_re);
    const bool con = (qe <= o_qe) & (re <= o_re);
    return near & (ext | con);
}

// The wide list of one merge: entries below near_cap in `near` (LDS when the
// caller has it), the rest at the same indices of `far`. near_cap is a
// multiple of 64, so the 64 entries of a chunk are all on one side.
// The LDS part is typed as LDS so that its accesses compile to LDS
// instructions rather than flat ones.
typedef __attribute__((address_space(3))) WideNam LdsWideNam;

struct WideList {
    LdsWideNam *near;
    WideNam *far;
    uint32_t near_cap;
};

// a wide list wholly in `p` (LDS or global)
__device__ __forceinline__ WideList wide_in(WideNam *p) { return WideList{nullptr, p, 0}; }

__device__ __forceinline__ WideNam wget(const WideList &L, uint32_t e) {
    WideNam w;
    if (e < L.near_cap) {
        const LdsWideNam &x = L.near[e];
        w = WideNam{x.qs, x.qe, x.qp, x.rs, x.re, x.rp, x.n, x.cid};
    } else {
        w = L.far[e];
    }
    return w;
}

__device__ __forceinline__ void wput(const WideList &L, uint32_t e, const WideNam &w) {
    if (e < L.near_cap) {
        LdsWideNam &x = L.near[e];
        x.qs = w.qs;
        x.qe = w.qe;
        x.qp = w.qp;
        x.rs = w.rs;
        x.re = w.re;
        x.rp = w.rp;
        x.n = w.n;
        x.cid = w.cid;
    } else {
        L.far[e] = w;
    }
}

// write the wide list's NAMs of the group that ends (all of them are final)
__device__ __forceinline__ void wide_flush(const WideList &wide, uint32_t nw, gis_nam *out, uint32_t base,
                                           uint32_t grp, uint32_t cap) {
    for (uint32_t c0 = 0; c0 < nw; c0 += 64) {
        const uint32_t e = c0 + lane_id();
        if (e < nw) {
            const WideNam w = wget(wide, e);
            if (base + w.cid < cap) lane_nam_store(out, base + w.cid, grp, from_wide(w));
        }
    }
}

// Grouped order of nh > 64 hits: nullptr when they already are in (strand,
// ref) order, else keys (bitonic in keys, nh <= MERGE_LDS) or one of keys /
// keys_alt (stable radix sort) holding group << 32 | hit index.
__device__ __forceinline__ const uint64_t *group_order(uint32_t nh, const DevHit *hits, uint64_t *keys,
                                                      uint64_t *keys_alt) {
    const uint32_t lane = lane_id();
    const uint64_t *order = nullptr;
    uint32_t unsorted = 0, max_ref = 0;
    for (uint32_t i0 = 0; i0 < nh; i0 += 64) {
        const uint32_t i = i0 + lane;
        if (i < nh) {
            const uint32_t rsd = hits[i].ref_strand;
            max_ref = max_ref > (rsd & 0x7fffffffu) ? max_ref : (rsd & 0x7fffffffu);
            if (i > 0 && hits[i - 1].ref_strand > rsd) unsorted = 1;
        }
    }
    if (__ballot(unsorted != 0) != 0) {
        if (nh <= (uint32_t)MERGE_LDS) {
            uint32_t P = 1;
            while (P < nh) P <<= 1;
            for (uint32_t i = lane; i < P; i += 64)
                keys[i] = i < nh ? (((uint64_t)hits[i].ref_strand << 32) | i) : UINT64_MAX;
            wave_sync_mem();
            bitonic_mem(keys, P);
            order = keys;
        } else {
            // compact group key: strand above the bits of the largest ref_id
            const uint32_t mr = wave_max(max_ref);
            const uint32_t rbits = mr ? 32u - (uint32_t)__builtin_clz(mr) : 0u;
            for (uint32_t i = lane; i < nh; i += 64) {
                const uint32_t rsd = hits[i].ref_strand;
                const uint64_t g = ((uint64_t)(rsd >> 31) << rbits) | (rsd & 0x7fffffffu);
                keys[i] = (g << 32) | i;
            }
            wave_sync_mem();
            order = radix_sort_groups(keys, keys_alt, nh, rbits + 1);
        }
    }
    return order;
}

// Reference-interval bounds of one 64-entry chunk of the wide list: every
// entry's previous hit start is at least lo and its ref end at most hi. Starts only grow and ends are raised with the entry, so the bounds
// stay valid (if loose) until the list is compacted and they are recomputed.
struct ChunkBox {
    int lo, hi;
};

__device__ __forceinline__ int wave_min_i(int v) {
    for (int d = 32; d >= 1; d >>= 1) v = min(v, __shfl_xor(v, d, 64));
    return v;
}

__device__ __forceinline__ int wave_max_i(int v) {
    for (int d = 32; d >= 1; d >>= 1) v = max(v, __shfl_xor(v, d, 64));
    return v;
}

// bounds of the wide list's chunks c0 .. (all of them when c0 == 0)
__device__ __forceinline__ void boxes_rebuild(const WideList &wide, uint32_t nw, ChunkBox *box) {
    for (uint32_t c = 0; c * 64 < nw; c++) {
        const uint32_t e = c * 64 + lane_id();
        int lo = 0x7fffffff, hi = (int)0x80000000;
        if (e < nw) {
            const WideNam w = wget(wide, e);
            lo = w.rp;
            hi = w.re;
        }
        lo = wave_min_i(lo);
        hi = wave_max_i(hi);
        if (lane_id() == 0) box[c] = ChunkBox{lo, hi};
    }
}

// Merge hits order[xb .. xe) (hit indices in grouped order; order nullptr:
// the hits themselves in order), or, with pre, the xe <= 64 hits already in
// the chunk registers. NAM c of the k-th group of the range is written at
// out[(NAMs of the range's earlier groups) + c]; entries at index >= cap are
// not written. Returns the NAM count, or MERGE_OVERFLOW when a group needs
// more than 64 open NAMs and there is no wide list.
// box (may be nullptr): bounds of up to box_cap chunks of the wide list. A
// hit can only be accepted by an entry with previous hit start < ref start <=
// ref end, so the scan for the first accepting entry visits only the chunks
// whose bounds admit the hit, in order: the same entry as a full scan.
__device__ __forceinline__ uint32_t merge_range(const DevHit *hits, const uint64_t *order, uint32_t xb,
                                                uint32_t xe, bool pre, uint32_t c_rsd, uint32_t c_rs,
                                                uint32_t c_re, uint32_t c_q, const WideList wide, gis_nam *out,
                                                uint32_t cap, ChunkBox *box = nullptr, uint32_t box_cap = 0) {
    const uint32_t lane = lane_id();
    LaneNam o{0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t n_open = 0, grp = 0xffffffffu, base = 0, gcnt = 0;
    bool wide_mode = false;
    uint32_t nw = 0, wide_mark = 0;
    for (uint32_t x0 = xb; x0 < xe; x0 += 64) {
        if (!pre) {
            const uint32_t i = x0 + lane;
            if (i < xe) {
                const DevHit h = hits[order ? (uint32_t)order[i] : i];
                c_rsd = h.ref_strand;
                c_rs = (uint32_t)h.ref_start;
                c_re = (uint32_t)h.ref_end;
                c_q = (uint32_t)h.query_start | ((uint32_t)h.query_end << 16);
            }
        }
        const uint32_t m = xe - x0 < 64 ? xe - x0 : 64;
        uint32_t no_block_until = 0;  // the rest of a block that did not fit goes hit by hit
        for (uint32_t t = 0; t < m; t++) {
            const uint32_t rsd = readlane_u32(c_rsd, t);
            const int rs = (int)readlane_u32(c_rs, t);
            const int re = (int)readlane_u32(c_re, t);
            const uint32_t q = readlane_u32(c_q, t);
            const int qs = (int)(q & 0xffffu), qe = (int)(q >> 16);
            if (rsd != grp) {
                if (wide_mode) {
                    wide_flush(wide, nw, out, base, grp, cap);
                    wide_mode = false;
                    nw = 0;
                } else if (lane < n_open && base + o.cid < cap) {
                    lane_nam_store(out, base + o.cid, grp, o);
                }
                base += gcnt;
                gcnt = 0;
                n_open = 0;
                grp = rsd;
            }
            if (!wide_mode) {
                // A block of hits with this hit's group and query start: as
                // in the wide form below, no NAM can take two of them (a NAM
                // that takes one gets previous start qs, and so does one a
                // hit of the block opens), so each meets the open NAMs as
                // they were before the block. Their acceptance ballots are
                // independent; only the assignment, in hit order (first
                // accepting NAM not taken by an earlier hit of the block,
                // else the next new lane), runs hit by hit, on scalars.
                const uint64_t sm =
                    __ballot(lane >= t && lane < m && c_rsd == rsd && (c_q & 0xffffu) == (uint32_t)qs);
                const uint64_t above = sm >> t;
                const uint32_t len = ~above ? (uint32_t)__builtin_ctzll(~above) : 64u - t;
                uint32_t n_new = 0;
                int my_hit = -1;  // the block hit (chunk lane) this lane takes or opens with
                bool fits = len > 1 && t >= no_block_until;
                if (fits) {
                    uint64_t taken = 0;
                    for (uint32_t i = t; i < t + len; i++) {
                        const int hrs = (int)readlane_u32(c_rs, i);
                        const int hre = (int)readlane_u32(c_re, i);
                        const int hqe = (int)(readlane_u32(c_q, i) >> 16);
                        bool hext = false;
                        const bool hacc = lane < n_open && nam_accepts(o.qp, o.qe, o.rp, o.re, qs, hqe, hrs, hre, hext);
                        const uint64_t avail = __ballot(hacc) & ~taken;
                        if (!avail && n_open + n_new == 64) {
                            // a new NAM would not fit the lanes: this block goes hit by hit
                            fits = false;
                            no_block_until = t + len;
                            break;
                        }
                        const uint32_t win = avail ? (uint32_t)__builtin_ctzll(avail) : n_open + n_new;
                        taken |= avail ? (1ull << win) : 0ull;
                        n_new += avail ? 0u : 1u;
                        my_hit = lane == win ? (int)i : my_hit;
                    }
                }
                if (fits) {
                    const int src = my_hit < 0 ? 0 : my_hit;
                    const int hrs = __shfl((int)c_rs, src, 64), hre = __shfl((int)c_re, src, 64);
                    const int hqe = (int)((uint32_t)__shfl((int)c_q, src, 64) >> 16);
                    const bool take = my_hit >= 0 && lane < n_open;
                    const bool mk = my_hit >= 0 && lane >= n_open;
                    const bool grow = mk || (take && hqe > o.qe && hre > o.re);
                    o.qs = mk ? qs : o.qs;
                    o.rs = mk ? hrs : o.rs;
                    o.qe = grow ? hqe : o.qe;
                    o.re = grow ? hre : o.re;
                    o.qp = take || mk ? qs : o.qp;
                    o.rp = take || mk ? hrs : o.rp;
                    o.n = mk ? 1 : o.n + (take ? 1 : 0);
                    o.cid = mk ? gcnt + (lane - n_open) : o.cid;
                    n_open += n_new;
                    gcnt += n_new;
                    t += len - 1;
                    continue;
                }
                bool ext = false;
                const bool acc = lane < n_open && nam_accepts(o.qp, o.qe, o.rp, o.re, qs, qe, rs, re, ext);
                const uint64_t bal = __ballot(acc);
                if (bal || n_open < 64) {
                    // the first accepting NAM takes the hit, or, with none, the
                    // hit opens NAM n_open; predicated updates, no branches
                    const bool fresh = bal == 0;
                    const uint32_t win = fresh ? n_open : (uint32_t)__builtin_ctzll(bal);
                    const bool me = lane == win;
                    const bool end = me && (fresh || ext);
                    o.qs = me && fresh ? qs : o.qs;
                    o.rs = me && fresh ? rs : o.rs;
                    o.qe = end ? qe : o.qe;
                    o.re = end ? re : o.re;
                    o.qp = me ? qs : o.qp;
                    o.rp = me ? rs : o.rp;
                    o.n = me ? (fresh ? 1 : o.n + 1) : o.n;
                    o.cid = me && fresh ? gcnt : o.cid;
                    n_open += fresh ? 1u : 0u;
                    gcnt += fresh ? 1u : 0u;
                    continue;
                }
                if (n_open == 64) {
                    const bool keep = o.qe >= qs;
                    const uint64_t km = __ballot(keep);
                    const uint32_t nk = (uint32_t)__popcll(km);
                    if (nk == 64) {
                        if (!wide.near && !wide.far) return MERGE_OVERFLOW;
                        // the group continues on the wide list, in creation order
                        wput(wide, lane, WideNam{o.qs, o.qe, o.qp, o.rs, o.re, o.rp, (uint32_t)o.n, o.cid});
                        nw = 64;
                        wide_mark = 64;
                        wide_mode = true;
                        n_open = 0;
                        if (box && box_cap) {
                            const int lo = wave_min_i(o.rp), hi = wave_max_i(o.re);
                            if (lane == 0) box[0] = ChunkBox{lo, hi};
                        }
                    } else {
                        if (!keep && base + o.cid < cap) lane_nam_store(out, base + o.cid, grp, o);
                        const uint32_t dst = keep ? mbcnt(km) : nk + mbcnt(~km);
                        o.qs = permute_push(o.qs, dst);
                        o.qe = permute_push(o.qe, dst);
                        o.qp = permute_push(o.qp, dst);
                        o.rs = permute_push(o.rs, dst);
                        o.re = permute_push(o.re, dst);
                        o.rp = permute_push(o.rp, dst);
                        o.n = permute_push(o.n, dst);
                        o.cid = (uint32_t)permute_push((int)o.cid, dst);
                        n_open = nk;
                    }
                }
                if (!wide_mode) {
                    if (lane == n_open) o = LaneNam{qs, qe, qs, rs, re, rs, 1, gcnt};
                    n_open++;
                    gcnt++;
                    continue;
                }
                // (the hit was accepted by no open NAM: it opens one on the wide list)
            } else {
                // Wide form, a block at a time: hit t and the hits after it in
                // the chunk registers with the same group and query start. A
                // NAM that takes a hit of the block gets previous start qs and
                // can take no other, and a NAM opened by one cannot either, so
                // every hit of the block meets the list as it was before the
                // block. Hit by hit, each takes the first accepting entry that
                // has not taken one of the block's hits. Going through the
                // list's chunks in order and, inside a chunk, through the
                // block's pending hits in order gives each hit that same entry
                // (the hits a chunk leaves pending find no acceptor in it, by
                // induction over the chunks), and a chunk is read once per
                // block instead of once per hit.
                const uint64_t sm =
                    __ballot(lane >= t && lane < m && c_rsd == rsd && (c_q & 0xffffu) == (uint32_t)qs);
                const uint64_t above = sm >> t;
                const uint32_t len = ~above ? (uint32_t)__builtin_ctzll(~above) : 64u;
                uint64_t pend = (len >= 64 ? ~0ull : ((1ull << len) - 1)) << t;
                const uint32_t nch = (nw + 63) >> 6;
                const bool boxed = box && nch <= box_cap;
                const int my_rs = (int)c_rs;
                for (uint32_t c = 0; c < nch && pend; c++) {
                    // the pending hits the chunk's bounds admit (all of them
                    // without bounds); the bounds of the chunk's start hold
                    // for every entry that has not taken a hit of the block
                    uint64_t cm = pend;
                    if (boxed) {
                        const ChunkBox bx = box[c];
                        cm = __ballot(((pend >> lane) & 1ull) != 0 && bx.lo < my_rs && my_rs <= bx.hi);
                    }
                    if (!cm) continue;
                    const uint32_t e = c * 64 + lane;
                    const bool valid = e < nw;
                    WideNam w{};
                    if (valid) w = wget(wide, e);
                    // the admitted hits, in order, take the first accepting
                    // entry of the chunk that no earlier one took: the
                    // ballots against the chunk as it was (an entry that
                    // takes a hit of the block can take no other), the
                    // assignment on scalars, the updates after
                    uint64_t taken = 0;
                    int my_hit = -1;  // the hit (chunk lane) this entry takes
                    while (cm) {
                        const uint32_t i = (uint32_t)__builtin_ctzll(cm);
                        cm &= cm - 1;
                        const int hrs = (int)readlane_u32(c_rs, i);
                        const int hre = (int)readlane_u32(c_re, i);
                        const int hqe = (int)(readlane_u32(c_q, i) >> 16);
                        bool ext = false;
                        const bool acc = valid && nam_accepts(w.qp, w.qe, w.rp, w.re, qs, hqe, hrs, hre, ext);
                        const uint64_t bal = __ballot(acc) & ~taken;
                        if (bal) {
                            const uint32_t win = (uint32_t)__builtin_ctzll(bal);
                            taken |= 1ull << win;
                            my_hit = lane == win ? (int)i : my_hit;
                            pend &= ~(1ull << i);
                        }
                    }
                    if (taken) {
                        const int src = my_hit < 0 ? 0 : my_hit;
                        const int hrs = __shfl((int)c_rs, src, 64), hre = __shfl((int)c_re, src, 64);
                        const int hqe = (int)((uint32_t)__shfl((int)c_q, src, 64) >> 16);
                        const bool tk = my_hit >= 0;
                        const bool ext = tk && hqe > w.qe && hre > w.re;
                        if (ext) {
                            w.qe = hqe;
                            w.re = hre;
                        }
                        if (tk) {
                            w.qp = qs;
                            w.rp = hrs;
                            w.n++;
                        }
                        if (boxed) {
                            const int hi = wave_max_i(ext ? hre : (int)0x80000000);
                            if (lane == 0 && box[c].hi < hi) box[c].hi = hi;
                        }
                        if (tk) wput(wide, e, w);
                        wave_sync_mem();
                    }
                }
                // the hits no entry took open entries, in hit order (lane 0
                // writes them; the wave syncs before a retirement reads the
                // list and after the last one)
                const bool opened = pend != 0;
                while (pend) {
                    const uint32_t i = (uint32_t)__builtin_ctzll(pend);
                    pend &= pend - 1;
                    const int hrs = (int)readlane_u32(c_rs, i);
                    const int hre = (int)readlane_u32(c_re, i);
                    const int hqe = (int)(readlane_u32(c_q, i) >> 16);
                    if (nw >= 2 * wide_mark) {
                        wave_sync_mem();
                        // retire the dead entries (query_end before this hit), keeping the order
                        uint32_t kept = 0;
                        for (uint32_t c0 = 0; c0 < nw; c0 += 64) {
                            const uint32_t e = c0 + lane;
                            WideNam w{};
                            bool keep = false;
                            if (e < nw) {
                                w = wget(wide, e);
                                keep = w.qe >= qs;
                                if (!keep && base + w.cid < cap)
                                    lane_nam_store(out, base + w.cid, grp, from_wide(w));
                            }
                            const uint64_t kb = __ballot(keep);
                            wave_sync_mem();
                            if (keep) wput(wide, kept + mbcnt(kb), w);
                            kept += (uint32_t)__popcll(kb);
                            wave_sync_mem();
                        }
                        nw = kept;
                        wide_mark = kept > 64 ? kept : 64;
                        if (box && ((nw + 63) >> 6) <= box_cap) boxes_rebuild(wide, nw, box);
                    }
                    if (lane == 0) {
                        wput(wide, nw, WideNam{qs, hqe, qs, hrs, hre, hrs, 1u, gcnt});
                        const uint32_t c = nw >> 6;
                        if (box && c < box_cap && ((nw + 63) >> 6) <= box_cap) {
                            if ((nw & 63) == 0) {
                                box[c] = ChunkBox{hrs, hre};
                            } else {
                                if (box[c].lo > hrs) box[c].lo = hrs;
                                if (box[c].hi < hre) box[c].hi = hre;
                            }
                        }
                    }
                    nw++;
                    gcnt++;
                }
                if (opened) wave_sync_mem();
                t += len - 1;
                continue;
            }
            if (lane == 0) {
                wput(wide, nw, WideNam{qs, qe, qs, rs, re, rs, 1u, gcnt});
                // the new entry's chunk bounds (only while every chunk has them)
                const uint32_t c = nw >> 6;
                if (box && c < box_cap && ((nw + 63) >> 6) <= box_cap) {
                    if ((nw & 63) == 0) {
                        box[c] = ChunkBox{rs, re};
                    } else {
                        if (box[c].lo > rs) box[c].lo = rs;
                        if (box[c].hi < re) box[c].hi = re;
                    }
                }
            }
            nw++;
            gcnt++;
            wave_sync_mem();
        }
    }
    if (wide_mode) wide_flush(wide, nw, out, base, grp, cap);
    else if (lane < n_open && base + o.cid < cap) lane_nam_store(out, base + o.cid, grp, o);
    return base + gcnt;
}

// Merge the nh hits of a read into NAMs written to out[0 .. total) in
// canonical order (entries at index >= cap are not written). keys: scratch
// of the next power of two >= nh entries when nh <= MERGE_LDS (bitonic
// grouping), else of nh entries with keys_alt another nh (radix grouping).
// wide: a list of nh entries for groups with more than 64 open NAMs, or
// nullptr (then such a read returns MERGE_OVERFLOW).
__device__ __forceinline__ uint32_t merge_wave(uint32_t nh, const DevHit *hits, uint64_t *keys,
                                               uint64_t *keys_alt, WideNam *wide, gis_nam *out,
                                               uint32_t cap) {
    const uint32_t lane = lane_id();
    if (nh > 64) return merge_range(hits, group_order(nh, hits, keys, keys_alt), 0, nh, false, 0, 0, 0, 0,
                                    wide_in(wide), out, cap);
    DevHit h{};
    if (lane < nh) h = hits[lane];
    const uint32_t rsd = lane < nh ? h.ref_strand : 0xffffffffu;
    const uint32_t prev = __shfl_up(rsd, 1, 64);
    const bool uns = __ballot(lane > 0 && lane < nh && prev > rsd) != 0;
    uint32_t c_rsd = rsd, c_rs = (uint32_t)h.ref_start, c_re = (uint32_t)h.ref_end;
    uint32_t c_q = (uint32_t)h.query_start | ((uint32_t)h.query_end << 16);
    if (uns) {
        uint64_t key = lane < nh ? (((uint64_t)rsd << 32) | lane) : UINT64_MAX;
        bitonic_regs(key);
        const int src = (int)((uint32_t)key & 63u);
        c_rsd = (uint32_t)__shfl((int)c_rsd, src, 64);
        c_rs = (uint32_t)__shfl((int)c_rs, src, 64);
        c_re = (uint32_t)__shfl((int)c_re, src, 64);
        c_q = (uint32_t)__shfl((int)c_q, src, 64);
    }
    return merge_range(hits, nullptr, 0, nh, true, c_rsd, c_rs, c_re, c_q, wide_in(wide), out, cap);
}

struct alignas(8) MergeLds {
    DevHit hits[MERGE_LDS];
    uint64_t keys[MERGE_LDS];
};

// reads handed from k_merge4 to k_merge (more than 64 hits, or more than
// 16 NAMs on one strand)
constexpr uint32_t MERGE_DEFER = 0xfffffffeu;
// reads whose (strand, ref) groups merge on separate waves (k_merge_groups,
// then k_merge_split_finish), for reads with more than SPLIT_MIN hits
constexpr uint32_t MERGE_SPLIT = 0xfffffffdu;
constexpr uint32_t SPLIT_MIN = 512;
// hits of one k_merge_groups work item: consecutive small groups of a split
// read merge on one wave
constexpr uint32_t ITEM_HITS = 384;
// wide lists in LDS: per wave of k_merge for reads of at most this many
// hits, and per group of k_merge_groups (one wave per block)
constexpr int MERGE_WIDE_LDS = 256;
constexpr int GROUP_WIDE_LDS = 1024;
// chunk bounds of a k_merge_groups wide list (lists up to 32,768 entries)
constexpr uint32_t GROUP_BOXES = 512;

__global__ __launch_bounds__(64 * WPB) void k_merge(DevBatch b, ReadSet set, int rescue, int only_deferred,
                                                    int split) {
    __shared__ uint64_t lkeys[WPB][MERGE_LDS];
    __shared__ WideNam lwide[WPB][MERGE_WIDE_LDS];  // the wide list of reads with few enough hits
    const uint32_t lane = lane_id();
    const int wid = (int)uni32(threadIdx.x >> 6);
    if (rescue ? over_capacity(b.rhit_off, b.n_reads, b.rhit_cap, b.ctl, OVF_RHITS)
               : over_capacity(b.hit_off, b.n_reads, b.hit_cap, b.ctl, OVF_HITS))
        return;
    const uint32_t cnt = set_count(set);
    for (uint32_t w = wave_index(WPB); w < cnt; w += gridDim.x * WPB) {
        const uint32_t r = uni32(set_at(set, w));
        const bool rescued = (uni32(b.flags[r]) & GIS_FLAG_RESCUED) != 0;
        if (rescued != (rescue != 0)) continue;
        if (only_deferred && uni32(b.nam_cnt[r]) != MERGE_DEFER) continue;
        const uint64_t hb = uni64(rescue ? b.rhit_off[r] : b.hit_off[r]);
        const uint32_t nh = uni32(rescue ? b.rhit_total[r] : b.hit_total[r]);
        if (nh == 0) {
            if (lane == 0) b.nam_cnt[r] = 0;
            continue;
        }
        const DevHit *ghits = (rescue ? b.rhits : b.hits) + hb;
        gis_nam *fin = (rescue ? b.rnam_fin : b.nam_fin) + hb;
        uint64_t *gkeys = b.merge_scratch ? (rescue ? b.rkeys : b.keys) + 2 * hb : nullptr;
        WideNam *gwide = b.merge_scratch ? (rescue ? b.rnam_ws : b.nam_ws) + hb : nullptr;
        if (split && nh > SPLIT_MIN && b.merge_scratch) {
            // group the hits; each group merges on its own wave
            const uint64_t *ord = group_order(nh, ghits, gkeys, gkeys + nh);
            const uint32_t code = ord == nullptr ? 0u : (ord == gkeys ? 1u : 2u);
            uint64_t *slots = code == 1u ? gkeys + nh : gkeys;
            uint32_t G = 0;
            for (uint32_t i0 = 0; i0 < nh; i0 += 64) {
                const uint32_t i = i0 + lane;
                bool head = false;
                if (i < nh) {
                    const uint32_t cur = ghits[ord ? (uint32_t)ord[i] : i].ref_strand;
                    head = i == 0 || ghits[ord ? (uint32_t)ord[i - 1] : i - 1].ref_strand != cur;
                }
                const uint64_t bal = __ballot(head);
                if (head) slots[G + mbcnt(bal)] = (uint64_t)i << 32;
                G += (uint32_t)__popcll(bal);
            }
            wave_sync_mem();
            // work items: runs of consecutive groups of at least ITEM_HITS
            // hits (a larger group is an item of its own); each item's start
            // goes to slots[item], which later holds its NAM count
            uint32_t K = 0;
            if (lane == 0) {
                uint32_t cur = 0;
                for (uint32_t g = 1; g < G; g++) {
                    const uint32_t st = (uint32_t)(slots[g] >> 32);
                    if (st - cur >= ITEM_HITS) {
                        slots[K++] = (uint64_t)cur << 32;  // K < g: heads not read yet stay intact
                        cur = st;
                    }
                }
                slots[K++] = (uint64_t)cur << 32;
            }
            K = (uint32_t)__builtin_amdgcn_readfirstlane((int)K);
            wave_sync_mem();
            {
                uint32_t at = 0;
                if (lane == 0) at = atomicAdd(&b.ctl->n_gwork[rescue ? 1 : 0], K);
                at = (uint32_t)__builtin_amdgcn_readfirstlane((int)at);
                for (uint32_t g = lane; g < K; g += 64) {
                    const uint32_t gs = (uint32_t)(slots[g] >> 32);
                    const uint32_t ge = g + 1 < K ? (uint32_t)(slots[g + 1] >> 32) : nh;
                    b.gwork[at + g] = make_uint4(r, gs, ge, g | (code << 30));
                }
                if (lane == 0) {
                    const uint32_t si = atomicAdd(&b.ctl->n_split[rescue ? 1 : 0], 1u);
                    b.split_list[si] = make_uint4(r, K, code, 0u);
                    b.nam_cnt[r] = MERGE_SPLIT;
                }
                wave_sync_mem();
                continue;
            }
        }
        uint32_t total = MERGE_OVERFLOW;
        if (nh <= (uint32_t)MERGE_WIDE_LDS) gwide = lwide[wid];  // NAMs <= hits fit
        if (nh <= (uint32_t)MERGE_LDS)
            total = merge_wave(nh, ghits, lkeys[wid], nullptr, gwide, fin, nh);
        else if (b.merge_scratch)
            total = merge_wave(nh, ghits, gkeys, gkeys + nh, gwide, fin, nh);
        if (total == MERGE_OVERFLOW) {
            // needs the global scratch, which this run does not have yet
            if (lane == 0) {
                atomicOr(&b.ctl->overflow, OVF_MERGE);
                b.nam_cnt[r] = 0;
            }
            continue;
        }
        if (lane == 0) b.nam_cnt[r] = total;
        wave_sync_mem();
    }
}

// one wave per work item (a run of consecutive groups of a split read): its
// hits (in grouped order) merge into NAMs at the item's own hit-indexed region
// of the read's output; the NAM count goes to the read's item slot
__global__ __launch_bounds__(64) void k_merge_groups(DevBatch b, int rescue) {
    __shared__ WideNam lwide[GROUP_WIDE_LDS];
    __shared__ ChunkBox lbox[GROUP_BOXES];
    const uint32_t lane = lane_id();
    const unsigned int cnt = b.ctl->n_gwork[rescue ? 1 : 0];
    for (uint32_t w = blockIdx.x; w < cnt; w += gridDim.x) {
        const uint4 it = b.gwork[w];
        const uint32_t r = it.x, gs = it.y, ge = it.z, g = it.w & 0x3fffffffu, code = it.w >> 30;
        const uint64_t hb = rescue ? b.rhit_off[r] : b.hit_off[r];
        const uint32_t nh = rescue ? b.rhit_total[r] : b.hit_total[r];
        uint64_t *gkeys = (rescue ? b.rkeys : b.keys) + 2 * hb;
        const uint64_t *ord = code == 0u ? nullptr : (code == 1u ? gkeys : gkeys + nh);
        uint64_t *slots = code == 1u ? gkeys + nh : gkeys;
        const DevHit *ghits = (rescue ? b.rhits : b.hits) + hb;
        // the first GROUP_WIDE_LDS wide entries in LDS, the rest in the
        // group's global region (as many entries as hits)
        WideNam *far = (rescue ? b.rnam_ws : b.nam_ws) + hb + gs;
        gis_nam *out = (rescue ? b.rnam_fin : b.nam_fin) + hb + gs;
        const WideList wl{(LdsWideNam *)lwide, far, (uint32_t)GROUP_WIDE_LDS};
        const uint32_t c = merge_range(ghits, ord, gs, ge, false, 0, 0, 0, 0, wl, out, ge - gs, lbox, GROUP_BOXES);
        if (lane == 0) slots[g] = ((uint64_t)gs << 32) | c;
        wave_sync_mem();
    }
}

// one wave per split read: the groups' NAMs move to consecutive places in
// group order (canonical order) with nam_id = rank
__global__ __launch_bounds__(64 * WPB) void k_merge_split_finish(DevBatch b, int rescue) {
    const uint32_t lane = lane_id();
    const unsigned int cnt = b.ctl->n_split[rescue ? 1 : 0];
    for (uint32_t w = blockIdx.x * WPB + (threadIdx.x >> 6); w < cnt; w += gridDim.x * WPB) {
        const uint4 e = b.split_list[w];
        const uint32_t r = e.x, G = e.y, code = e.z;
        const uint64_t hb = rescue ? b.rhit_off[r] : b.hit_off[r];
        const uint32_t nh = rescue ? b.rhit_total[r] : b.hit_total[r];
        uint64_t *gkeys = (rescue ? b.rkeys : b.keys) + 2 * hb;
        const uint64_t *slots = code == 1u ? gkeys + nh : gkeys;
        gis_nam *fin = (rescue ? b.rnam_fin : b.nam_fin) + hb;
        uint32_t base = 0;
        for (uint32_t g = 0; g < G; g++) {
            const uint64_t sg = slots[g];
            const uint32_t gs = (uint32_t)(sg >> 32), c = (uint32_t)sg;
            // base <= gs: every NAM moves towards the front, chunk by chunk
            for (uint32_t c0 = 0; c0 < c; c0 += 64) {
                const uint32_t k = c0 + lane;
                gis_nam x{};
                if (k < c) x = fin[gs + k];
                wave_sync_mem();
                if (k < c) {
                    x.nam_id = (int32_t)(base + k);
                    fin[base + k] = x;
                }
                wave_sync_mem();
            }
            base += c;
        }
        if (lane == 0) b.nam_cnt[r] = base;
        wave_sync_mem();
    }
}

// ---------------------------------------------------------------------------
// Merge of four reads per wave (16 lanes each) for reads with at most 64
// hits. Each lane holds one NAM of its read; NAMs of every (strand, ref)
// group of the current strand share the segment, so the hits are taken in
// arrival order without grouping them: within one group that is
// (query_start, ref_start) order, and a NAM only accepts hits of its own
// group. Each segment broadcasts its read's next hit from LDS, a ballot
// restricted to the segment picks the first accepting NAM (creation order),
// and a hit nobody accepts opens a NAM in the segment's next lane. When a
// strand's hits end, its NAMs are final: each lane ranks its NAM among the
// segment's by (ref_id, creation index), which is canonical order within the
// strand, and writes it. A read with more than 16 NAMs on one strand, or
// more than 64 hits, is marked MERGE_DEFER for k_merge.

constexpr uint32_t SEG = 16;
constexpr uint32_t RPW = 64 / SEG;

__device__ __forceinline__ void seg_finalize(gis_nam *out, uint32_t base, uint32_t j, uint32_t n_open,
                                             uint32_t rsd, const LaneNam &o, bool on) {
    // rank by (ref_id, creation index) among the segment's open lanes. A
    // segment is one DPP row of 16 lanes: rotating the row by 1..15 shows
    // each lane every other lane's key. Creation indices are below 16, so
    // no key is all ones, which marks the lanes beyond n_open.
    const uint32_t key = ((rsd & 0x7fffffffu) << 8) | (o.cid & 0xffu);
    const int kv = (int)(j < n_open ? key : 0xffffffffu);
    uint32_t rank = 0;
#define GIS_ROW_ROR(r) rank += (uint32_t)__builtin_amdgcn_update_dpp(0, kv, 0x120 + (r), 0xf, 0xf, false) < key ? 1u : 0u
    GIS_ROW_ROR(1); GIS_ROW_ROR(2); GIS_ROW_ROR(3); GIS_ROW_ROR(4); GIS_ROW_ROR(5);
    GIS_ROW_ROR(6); GIS_ROW_ROR(7); GIS_ROW_ROR(8); GIS_ROW_ROR(9); GIS_ROW_ROR(10);
    GIS_ROW_ROR(11); GIS_ROW_ROR(12); GIS_ROW_ROR(13); GIS_ROW_ROR(14); GIS_ROW_ROR(15);
#undef GIS_ROW_ROR
    if (on && j < n_open) lane_nam_store(out, base + rank, rsd, o);
}

__global__ __launch_bounds__(64 * WPB) void k_merge4(DevBatch b, ReadSet set, int rescue) {
    __shared__ DevHit lh[WPB][RPW][64];
    const uint32_t lane = lane_id(), seg = lane / SEG, j = lane % SEG;
    const int wid = (int)uni32(threadIdx.x >> 6);
    if (rescue ? over_capacity(b.rhit_off, b.n_reads, b.rhit_cap, b.ctl, OVF_RHITS)
               : over_capacity(b.hit_off, b.n_reads, b.hit_cap, b.ctl, OVF_HITS))
        return;
    const uint32_t cnt = set_count(set);
    const uint32_t stride = gridDim.x * WPB * RPW;
    for (uint32_t w0 = wave_index(WPB) * RPW; w0 < cnt; w0 += stride) {
        const uint32_t widx = w0 + seg;
        bool act = widx < cnt;
        uint32_t r = 0, nh = 0;
        uint64_t hb = 0;
        if (act) {
            r = set_at(set, widx);
            const bool rescued = (b.flags[r] & GIS_FLAG_RESCUED) != 0;
            act = rescued == (rescue != 0);
        }
        if (act) {
            hb = rescue ? b.rhit_off[r] : b.hit_off[r];
            nh = rescue ? b.rhit_total[r] : b.hit_total[r];
            if (nh == 0 || nh > 64) {
                if (j == 0) b.nam_cnt[r] = nh == 0 ? 0u : MERGE_DEFER;
                act = false;
            }
        }
        if (act) {
            const DevHit *gh = (rescue ? b.rhits : b.hits) + hb;
            for (uint32_t k = j; k < nh; k += SEG) lh[wid][seg][k] = gh[k];
        }
        wave_sync_mem();
        gis_nam *out = (rescue ? b.rnam_fin : b.nam_fin) + hb;
        LaneNam o{0, 0, 0, 0, 0, 0, 0, 0};
        uint32_t rsd_o = 0;        // this lane's NAM group
        uint32_t n_open = 0;       // NAMs of the segment on the current strand
        uint32_t strand = 0, base = 0;
        bool ovf = false;
        const uint32_t maxh = wave_max_dpp(act ? nh : 0u);
        for (uint32_t t = 0; t < maxh; t++) {
            const bool valid = act && !ovf && t < nh;
            DevHit h{};
            if (valid) h = lh[wid][seg][t];
            const uint32_t hs = h.ref_strand >> 31;
            // the strand changes: the segment's forward NAMs are final
            const bool sw = valid && hs != strand;
            if (__ballot(sw)) {
                seg_finalize(out, base, j, n_open, rsd_o, o, sw);
                if (sw) {
                    base += n_open;
                    n_open = 0;
                    strand = hs;
                }
            }
            const int qs = h.query_start, qe = h.query_end, rs = h.ref_start, re = h.ref_end;
            bool ext = false;
            const bool acc = valid & (j < n_open) & (rsd_o == h.ref_strand) &
                             nam_accepts(o.qp, o.qe, o.rp, o.re, qs, qe, rs, re, ext);
            const uint64_t bal = __ballot(acc);
            const uint32_t mine = (uint32_t)(bal >> (seg * SEG)) & ((1u << SEG) - 1u);
            // the first accepting lane of the segment takes the hit; else the
            // hit opens a NAM in the segment's next lane (selects, no branches)
            const bool take = (mine != 0u) & (j == (uint32_t)__builtin_ctz(mine | 0x10000u));
            const bool opens = valid & (mine == 0u);
            const bool full = n_open == SEG;
            const bool mk = opens & !full & (j == n_open);
            o.qe = (take & ext) ? qe : o.qe;
            o.re = (take & ext) ? re : o.re;
            o.qp = (take | mk) ? qs : o.qp;
            o.rp = (take | mk) ? rs : o.rp;
            o.n = mk ? 1 : o.n + (take ? 1 : 0);
            o.qs = mk ? qs : o.qs;
            o.rs = mk ? rs : o.rs;
            o.qe = mk ? qe : o.qe;
            o.re = mk ? re : o.re;
            o.cid = mk ? n_open : o.cid;
            rsd_o = mk ? h.ref_strand : rsd_o;
            ovf = ovf | (opens & full);
            n_open += (opens & !full) ? 1u : 0u;
        }
        const bool fin = act && !ovf;
        if (__ballot(fin)) seg_finalize(out, base, j, n_open, rsd_o, o, fin);
        if (act && j == 0) b.nam_cnt[r] = ovf ? MERGE_DEFER : base + n_open;
        wave_sync_mem();
    }
}

// ---------------------------------------------------------------------------
// fused: the whole path of one read in one wave. Each lane holds randstrobe
// pair i (forward i, reverse complement i) in registers, so a read fits when
// it has at most 64 + w_min syncmers; hits and the merge live in LDS. Reads
// beyond these limits are appended to the spill list for the per-stage
// kernels.
//
// The kernel is persistent over a device work queue: a list of read ids
// ordered by length (longest first), cut by the host into a long and a short
// bin that run as separate launches with their own LDS tiling (LMAX). Waves
// take reads either statically (wave w takes queue entries w, w + W, ...,
// which deals the length-sorted queue round robin) or dynamically (a wave
// grabs `chunk` entries at a time from a counter in the control block).

constexpr int F_SYCAP = 96;

template <int LMAX>
struct alignas(8) FusedSeedLds {
    uint64_t sh[LMAX];
    uint64_t nmask[LMAX / 64 + 2];
    uint64_t sy_hash[F_SYCAP];
    uint64_t pk[2 * ((LMAX + 63) / 64) + 1];
    uint16_t sy_pos[F_SYCAP];
    uint8_t st[LMAX];
};

template <int LMAX>
union alignas(8) FusedLds {
    FusedSeedLds<LMAX> seed;
    MergeLds merge;
};

// hand read r to the per-stage kernels (reads spill rarely, so one atomic
// per spilled read does not contend)
__device__ __forceinline__ void spill_read(const DevBatch &b, uint32_t r) {
    if (lane_id() == 0) {
        b.flags[r] = FLAG_SPILL;
        b.nam_cnt[r] = 0;
        b.hit_total[r] = 0;
        b.rhit_total[r] = 0;
        const unsigned int at = atomicAdd(&b.ctl->n_spill, 1u);
        b.spill_ids[at] = r;
    }
    wave_sync_mem();
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// The bucket line of one query randstrobe, fetched in one round trip. The
// empty asm takes every loaded word, so all eight loads issue before the
// first wait: otherwise the compiler sinks the record loads into the
// small-bucket branch, behind the wait for the bucket's bounds, and each
// probe pays two dependent fetches. Each lane holds one query randstrobe,
// so a read's forward and reverse-complement lines arrive together.
__device__ __forceinline__ void probe_line_now(const DevIndex &ix, Probe &P) {
    const u32x4 *l = reinterpret_cast<const u32x4 *>(ix.lines) +
                     (uint64_t)(uint32_t)(P.key >> (64 - ix.bits)) * BUCKET_LINE;
    u32x4 a[BUCKET_LINE];
#pragma unroll
    for (uint32_t j = 0; j < BUCKET_LINE; j++) a[j] = l[j];
    asm volatile("" ::"v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(a[3]), "v"(a[4]), "v"(a[5]), "v"(a[6]), "v"(a[7]));
    P.s = (uint64_t)a[0].x | ((uint64_t)a[0].y << 32);
    P.e = P.s + a[0].z;
#pragma unroll
    for (uint32_t j = 0; j < PROBE_SMALL; j++)
        P.r[j] = IdxRec{(uint64_t)a[j + 1].x | ((uint64_t)a[j + 1].y << 32), a[j + 1].z, a[j + 1].w};
}

// find + is_filtered + the counting run walk of one query randstrobe through
// its fetched bucket line (the per-stage lookup's probe): position, found
// and not filtered, kept hits, and for a small bucket the kept records' mask
__device__ __forceinline__ void fused_probe(const DevIndex &ix, Probe &P, int qs, int qe, uint64_t &pos,
                                            bool &good, uint32_t &kept, uint32_t &mask, bool &small) {
    probe_finish(ix, P, qs, qe);
    pos = P.pos;
    good = P.found && !P.filtered;
    kept = P.hc;
    mask = P.keep;
    small = probe_small(P);
}

// the records of a small bucket's line that `mask` selects, written as hits
// in run order (the line is read again; it was fetched by the probe)
__device__ __forceinline__ void fused_emit_line(const DevIndex &ix, uint64_t key, uint32_t mask, int qs, int qe,
                                                uint32_t strand_bit, DevHit *out) {
    Probe P;
    P.key = key;
    probe_line(ix, P);
    P.keep = mask;
    probe_emit(ix, P, qs, qe, strand_bit, out);
}

// The lanes' runs flagged in `mine` walked by the whole wave, one after the
// other: the kept count goes to the owning lane's `kept`, and with out the
// hits go to out + (the owning lane's at), with the owning lane's strand bit.
__device__ __forceinline__ void fused_wave_walks(const DevIndex &ix, bool mine, uint64_t pos, uint64_t key,
                                                 uint32_t qs, uint32_t qe, uint32_t strand_bit, uint32_t &kept,
                                                 DevHit *out, uint32_t at) {
    const uint32_t lane = lane_id();
    uint64_t m = __ballot(mine);
    while (m) {
        const uint32_t l = (uint32_t)__builtin_ctzll(m);
        m &= m - 1;
        const uint64_t pl =
            (uint64_t)readlane_u32((uint32_t)pos, l) | ((uint64_t)readlane_u32((uint32_t)(pos >> 32), l) << 32);
        const uint64_t kl =
            (uint64_t)readlane_u32((uint32_t)key, l) | ((uint64_t)readlane_u32((uint32_t)(key >> 32), l) << 32);
        const uint32_t k = walk_run_wave(ix, pl, kl, (int)readlane_u32(qs, l), (int)readlane_u32(qe, l),
                                         readlane_u32(strand_bit, l), out ? out + readlane_u32(at, l) : nullptr);
        if (lane == l) kept = k;
    }
}

// One read through seeding, lookup and its hits. With `stash` (an LDS slot
// of 64 hits), a read with 1 to 64 hits leaves them there and reports their
// count in stashed for a merge with three other reads (fused_quad_merge);
// every other read is merged alone (or spilled) here.
template <int LMAX>
__device__ __forceinline__ void fused_read(const DevBatch &b, const DevIndex &ix, const SeedParams &P,
                                           FusedLds<LMAX> &S, uint32_t r, uint32_t rescue_cutoff,
                                           uint32_t hard_cutoff, DevHit *stash, uint32_t &stashed) {
    stashed = 0;
    const uint32_t lane = lane_id();
    const int L = (int)sload(b.read_len + r);
    bool spill = L > LMAX;
    uint32_t nh = 0, n_fwd = 0;
    bool rescue = false;
    float frac = 1.0f;
    if (!spill) {
        SeedBufs B{S.seed.pk, S.seed.st, S.seed.sh, S.seed.sy_hash, S.seed.sy_pos, S.seed.nmask};
        const uint32_t n = seed_syncmers(b.seqs + sload(b.read_off + r), L, P, B, F_SYCAP, nullptr, nullptr);
        n_fwd = ((uint32_t)L >= P.w_max && n > P.w_min) ? n - P.w_min : 0;
        spill = n > (uint32_t)F_SYCAP || n_fwd > 64;
        // query randstrobe j (randstrobes_query order: the forward ones, then
        // the reverse-complement ones) on lane j, and randstrobe 64 + j, when
        // the read has more than 64, on lane j's second slot
        const uint32_t nq = 2 * n_fwd;
        const bool aa = !spill && lane < nq, ab = !spill && lane + 64 < nq;
        Rs1 qa{}, qb{};
        if (aa) qa = randstrobe_one(B, n, n_fwd, lane, L, P);
        if (ab) qb = randstrobe_one(B, n, n_fwd, lane + 64, L, P);
        wave_sync();  // the syncmer lists are dead from here; LDS becomes MergeLds
        if (!spill) {
            // each randstrobe through find / filter / run walk: the
            // bucket-line probe of the per-stage lookup; a small bucket's kept
            // records are remembered as a mask over its line. Second-slot
            // randstrobes (64 and up, n_fwd <= 64) are all reverse-complement.
            const uint32_t sba = lane < n_fwd ? 0u : 0x80000000u, sbb = 0x80000000u;
            uint64_t pa = GIS_NOT_FOUND, pb = GIS_NOT_FOUND;
            uint32_t ha = 0, hb = 0, ka = 0, kb = 0;
            bool ga = false, gb = false, sa = false, sb = false;
            // the second slot first: the first slot's probe then stays in
            // registers for the hits of a read that is not rescued
            if (ab) {
                Probe Q;
                Q.key = qb.h;
                probe_line_now(ix, Q);
                fused_probe(ix, Q, (int)qb.s, (int)qb.e, pb, gb, hb, kb, sb);
            }
            Probe QA;
            QA.key = qa.h;
            if (aa) {
                probe_line_now(ix, QA);
                fused_probe(ix, QA, (int)qa.s, (int)qa.e, pa, ga, ha, ka, sa);
            }
            // (the whole wave is active here: spill and rescue are wave-uniform)
            const uint32_t found = wave_count(pa != GIS_NOT_FOUND) + wave_count(pb != GIS_NOT_FOUND);
            const uint32_t good = wave_count(ga) + wave_count(gb);
            const uint32_t hits_norm = wave_sum_dpp(ha + hb);
            frac = found > 0 ? (float)good / (float)found : 1.0f;
            rescue = (double)frac < 0.7 || hits_norm == 0;
            if (!rescue) {
                // hits into LDS in query order (forward seeds, then reverse
                // ones): the first slot's straight from the lines its probes
                // fetched, a second slot's lines read again
                const uint32_t ia = wave_incl_scan_dpp(ha), ib = nq > 64 ? wave_incl_scan_dpp(hb) : 0u;
                const uint32_t tot_a = readlane_u32(ia, 63);
                nh = tot_a + readlane_u32(ib, 63);
                spill = nh > (uint32_t)MERGE_LDS;
                if (!spill) {
                    DevHit *dst = stash && nh <= 64 ? stash : S.merge.hits;
                    if (ha) probe_emit(ix, QA, (int)qa.s, (int)qa.e, sba, dst + (ia - ha));
                    if (hb) {
                        DevHit *ob = dst + tot_a + (ib - hb);
                        if (sb) fused_emit_line(ix, qb.h, kb, (int)qb.s, (int)qb.e, sbb, ob);
                        else walk_run(ix, pb, qb.h, (int)qb.s, (int)qb.e, sbb, ob);
                    }
                    wave_sync_mem();
                }
            } else {
                bool wa = false, wb = false;  // taken runs longer than WALK_WAVE_MIN: the wave walks them
                {
                    // run lengths, and the taken prefix of each strand: a seed's
                    // rank among its strand's seeds in (count, start, end) order
                    const uint32_t ca = aa && pa != GIS_NOT_FOUND ? dev_get_count(ix, pa) : 0u;
                    const uint32_t cb = ab && pb != GIS_NOT_FOUND ? dev_get_count(ix, pb) : 0u;
                    const bool fwd_a = lane < n_fwd;
                    uint32_t ranka = 0, rankb = 0;
                    for (uint32_t j = 0; j < n_fwd; j++) {
                        // forward randstrobe j (lane j) and reverse-complement
                        // randstrobe n_fwd + j (lane or second slot)
                        const uint32_t cf = readlane_u32(ca, j), sf = readlane_u32(qa.s, j), ef = readlane_u32(qa.e, j);
                        const uint32_t y = n_fwd + j;
                        const bool hi = y >= 64;
                        const uint32_t yl = hi ? y - 64 : y;
                        const uint32_t cr = hi ? readlane_u32(cb, yl) : readlane_u32(ca, yl);
                        const uint32_t sr = hi ? readlane_u32(qb.s, yl) : readlane_u32(qa.s, yl);
                        const uint32_t er = hi ? readlane_u32(qb.e, yl) : readlane_u32(qa.e, yl);
                        const uint32_t cj = fwd_a ? cf : cr, sj = fwd_a ? sf : sr, ej = fwd_a ? ef : er;
                        ranka += (cj > 0 && rh_less(cj, sj, ej, ca, qa.s, qa.e)) ? 1u : 0u;
                        rankb += (cr > 0 && rh_less(cr, sr, er, cb, qb.s, qb.e)) ? 1u : 0u;
                    }
                    const bool ta = ca > 0 && ca <= hard_cutoff && (ca <= rescue_cutoff || ranka < 5);
                    const bool tb = cb > 0 && cb <= hard_cutoff && (cb <= rescue_cutoff || rankb < 5);
                    wa = ta && ca > WALK_WAVE_MIN;
                    wb = tb && cb > WALK_WAVE_MIN;
                    ha = ta && !wa ? walk_run(ix, pa, qa.h, (int)qa.s, (int)qa.e, sba, nullptr) : 0u;
                    hb = tb && !wb ? walk_run(ix, pb, qb.h, (int)qb.s, (int)qb.e, sbb, nullptr) : 0u;
                    fused_wave_walks(ix, wa, pa, qa.h, qa.s, qa.e, sba, ha, nullptr, 0u);
                    fused_wave_walks(ix, wb, pb, qb.h, qb.s, qb.e, sbb, hb, nullptr, 0u);
                }
                // hits into LDS in query order: forward seeds, then reverse ones
                const uint32_t ia = wave_incl_scan_dpp(ha), ib = nq > 64 ? wave_incl_scan_dpp(hb) : 0u;
                const uint32_t tot_a = readlane_u32(ia, 63);
                nh = tot_a + readlane_u32(ib, 63);
                spill = nh > (uint32_t)MERGE_LDS;
                if (!spill) {
                    // a read that goes to the four-read merge writes its hits
                    // straight into its stash slot
                    DevHit *dst = stash && nh <= 64 ? stash : S.merge.hits;
                    const uint32_t oa = ia - ha, ob = tot_a + ib - hb;
                    // a taken seed of a small bucket keeps the records the probe
                    // kept (its run walk is the same in both modes)
                    if (ha && !wa) {
                        if (sa) fused_emit_line(ix, qa.h, ka, (int)qa.s, (int)qa.e, sba, dst + oa);
                        else walk_run(ix, pa, qa.h, (int)qa.s, (int)qa.e, sba, dst + oa);
                    }
                    if (hb && !wb) {
                        if (sb) fused_emit_line(ix, qb.h, kb, (int)qb.s, (int)qb.e, sbb, dst + ob);
                        else walk_run(ix, pb, qb.h, (int)qb.s, (int)qb.e, sbb, dst + ob);
                    }
                    uint32_t unused_a = 0, unused_b = 0;
                    fused_wave_walks(ix, wa, pa, qa.h, qa.s, qa.e, sba, unused_a, dst, oa);
                    fused_wave_walks(ix, wb, pb, qb.h, qb.s, qb.e, sbb, unused_b, dst, ob);
                    wave_sync_mem();
                }
            }
        }
    }
    if (spill) {
        spill_read(b, r);
        return;
    }
    if (stash && nh > 0 && nh <= 64) {
        // the hits are in the stash slot
        if (lane == 0) {
            b.frac[r] = frac;
            b.flags[r] = (uint8_t)((rescue ? GIS_FLAG_RESCUED : 0) | (n_fwd == 0 ? GIS_FLAG_NO_RANDSTROBES : 0));
            b.hit_total[r] = 0;
            b.rhit_total[r] = 0;
        }
        stashed = nh;
        return;
    }
    uint32_t total = 0;
    if (nh > 0)
        total = merge_wave(nh, S.merge.hits, S.merge.keys, nullptr, nullptr,
                           b.nam_slots + (uint64_t)r * F_NAMCAP, (uint32_t)F_NAMCAP);
    if (total > (uint32_t)F_NAMCAP) {  // includes MERGE_OVERFLOW
        spill_read(b, r);
        return;
    }
    if (lane == 0) {
        b.frac[r] = frac;
        b.flags[r] = (uint8_t)((rescue ? GIS_FLAG_RESCUED : 0) | (n_fwd == 0 ? GIS_FLAG_NO_RANDSTROBES : 0));
        b.nam_cnt[r] = total;
        b.hit_total[r] = 0;
        b.rhit_total[r] = 0;
    }
    wave_sync_mem();
}

// The merge of up to four stashed reads (read k's hits in stash[k] for
// k < 3, the fourth read's in slot3) on one wave, 16 lanes a read, as
// k_merge4 merges them. A read a strand of which needs more than 16 NAMs is
// then merged alone from its stash (merge_wave), and spilled if its NAMs
// exceed F_NAMCAP.
// meta: the reads' ids (0..3) and hit counts (4..7), in LDS.
__device__ __forceinline__ void fused_quad_merge(const DevBatch &b, DevHit (*stash)[64], DevHit *slot3, uint32_t qn,
                                                 const uint32_t *meta) {
    const uint32_t lane = lane_id(), seg = lane / SEG, j = lane % SEG;
    const bool act = seg < qn;
    const uint32_t r = meta[seg];
    const uint32_t nh = act ? meta[4 + seg] : 0u;
    const DevHit *mine = seg < 3 ? stash[seg] : slot3;
    gis_nam *out = b.nam_slots + (uint64_t)r * F_NAMCAP;
    LaneNam o{0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t rsd_o = 0, n_open = 0, strand = 0, base = 0;
    bool ovf = false;
    const uint32_t maxh = wave_max_dpp(nh);
    for (uint32_t t = 0; t < maxh; t++) {
        const bool valid = act && !ovf && t < nh;
        DevHit h{};
        if (valid) h = mine[t];
        const uint32_t hs = h.ref_strand >> 31;
        const bool sw = valid && hs != strand;
        if (__ballot(sw)) {
            seg_finalize(out, base, j, n_open, rsd_o, o, sw);
            if (sw) {
                base += n_open;
                n_open = 0;
                strand = hs;
            }
        }
        const int qs = h.query_start, qe = h.query_end, rs = h.ref_start, re = h.ref_end;
        bool ext = false;
        const bool acc = valid & (j < n_open) & (rsd_o == h.ref_strand) &
                         nam_accepts(o.qp, o.qe, o.rp, o.re, qs, qe, rs, re, ext);
        const uint64_t bal = __ballot(acc);
        const uint32_t mine = (uint32_t)(bal >> (seg * SEG)) & ((1u << SEG) - 1u);
        // as in k_merge4: the first accepting lane takes the hit, else it
        // opens a NAM in the segment's next lane (selects, no branches)
        const bool take = (mine != 0u) & (j == (uint32_t)__builtin_ctz(mine | 0x10000u));
        const bool opens = valid & (mine == 0u);
        const bool full = n_open == SEG;
        const bool mk = opens & !full & (j == n_open);
        o.qe = (take & ext) ? qe : o.qe;
        o.re = (take & ext) ? re : o.re;
        o.qp = (take | mk) ? qs : o.qp;
        o.rp = (take | mk) ? rs : o.rp;
        o.n = mk ? 1 : o.n + (take ? 1 : 0);
        o.qs = mk ? qs : o.qs;
        o.rs = mk ? rs : o.rs;
        o.qe = mk ? qe : o.qe;
        o.re = mk ? re : o.re;
        o.cid = mk ? n_open : o.cid;
        rsd_o = mk ? h.ref_strand : rsd_o;
        ovf = ovf | (opens & full);
        n_open += (opens & !full) ? 1u : 0u;
    }
    const bool fin = act && !ovf;
    if (__ballot(fin)) seg_finalize(out, base, j, n_open, rsd_o, o, fin);
    if (fin && j == 0) b.nam_cnt[r] = base + n_open;
    wave_sync_mem();
    uint64_t redo = __ballot(act && ovf && j == 0);
    while (redo) {
        const uint32_t l = (uint32_t)__builtin_ctzll(redo);
        redo &= redo - 1;
        const uint32_t rk = readlane_u32(r, l), nk = readlane_u32(nh, l);
        const uint32_t total =
            merge_wave(nk, l / SEG < 3 ? stash[l / SEG] : slot3, nullptr, nullptr, nullptr,
                       b.nam_slots + (uint64_t)rk * F_NAMCAP,
                       (uint32_t)F_NAMCAP);
        if (total > (uint32_t)F_NAMCAP) {
            spill_read(b, rk);
        } else {
            if (lane == 0) b.nam_cnt[rk] = total;
            wave_sync_mem();
        }
    }
}

// Six waves per SIMD: at most 80 VGPRs, and LDS for six 4-wave blocks per
// CU. A wave merges its reads four at a time (fused_quad_merge) once their
// hits wait in its stash: three slots in the stash, and the fourth read's
// hits left in the wave's own merge area, which nothing else touches until
// the four-read merge has run (the next read's seeding comes after it).
template <int LMAX, int FWPB, bool K20>
__global__ __launch_bounds__(64 * FWPB) __attribute__((amdgpu_waves_per_eu(6))) void k_fused(
    DevBatch b, DevIndex ix, SeedParams P, uint32_t rescue_cutoff, uint32_t hard_cutoff,
    const uint32_t *__restrict__ queue, uint32_t count, uint32_t chunk, unsigned int *head) {
    if (K20) pin_k20_profile(P);
    __shared__ FusedLds<LMAX> lds[FWPB];
    __shared__ DevHit qstash[FWPB][3][64];
    __shared__ uint32_t qmeta[FWPB][8];
    const int wid = (int)uni32(threadIdx.x >> 6);
    FusedLds<LMAX> &S = lds[wid];
    DevHit(*stash)[64] = qstash[wid];
    uint32_t *meta = qmeta[wid];
    uint32_t qn = 0;
    // a grid with a wave per read merges each read alone
    const bool alone = chunk == 0 && (uint64_t)gridDim.x * FWPB >= count;
    // static: queue entries w, w + W, ...; dynamic: grabs of `chunk` entries,
    // and every wave leaves once a grab starts at or past the end
    uint32_t w = chunk == 0 ? wave_index(FWPB) : 0u, end = 0;
    for (;;) {
        bool have;
        if (chunk == 0) {
            have = w < count;
        } else {
            if (w >= end) {
                uint32_t start = 0;
                if (lane_id() == 0) start = atomicAdd(head, chunk);
                start = (uint32_t)__builtin_amdgcn_readfirstlane((int)start);
                w = start;
                end = count - min(start, count) < chunk ? count : start + chunk;
            }
            have = w < end;
        }
        if (have) {
            const uint32_t r = sload(queue + w);
            w += chunk == 0 ? gridDim.x * FWPB : 1u;
            uint32_t nh = 0;
            DevHit *slot = qn < 3 ? stash[qn] : S.merge.hits;
            fused_read<LMAX>(b, ix, P, S, r, rescue_cutoff, hard_cutoff, alone ? nullptr : slot, nh);
            if (nh) {
                if (lane_id() == 0) {
                    meta[qn] = r;
                    meta[4 + qn] = nh;
                }
                wave_sync_mem();
                qn++;
            }
        }
        if (qn == 4 || (!have && qn)) {
            fused_quad_merge(b, stash, S.merge.hits, qn, meta);
            qn = 0;
        }
        if (!have) break;
    }
}

// ---------------------------------------------------------------------------
// dense output: NAMs of read r to nams_out[nam_off[r] ..]

constexpr uint32_t COMPACT_SEG = 16;
constexpr uint32_t COMPACT_SEG_WORDS = 8 * COMPACT_SEG;  // 11 NAMs of 11 words

__global__ __launch_bounds__(64 * WPB) void k_compact(DevBatch b, int fused) {
    const uint32_t lane = lane_id();
    if (blockIdx.x == 0 && threadIdx.x == 0) b.ctl->tot[2] = b.nam_off[b.n_reads];
    // a hit array overflowed: the merges did not run and the counts are stale
    if (b.ctl->overflow & (OVF_HITS | OVF_RHITS)) return;
    if (b.nam_off[b.n_reads] > b.out_cap) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&b.ctl->overflow, OVF_OUT);
        return;
    }
    // COMPACT_SEG lanes per read, so that a wave's reads start their loads
    // (count, flags, offsets, then NAMs) together; reads with more than
    // COMPACT_SEG_WORDS words of NAMs (repeats) are copied by the whole wave
    // afterwards, four words per lane in flight
    const uint32_t seg = lane / COMPACT_SEG, j = lane % COMPACT_SEG;
    const uint32_t per_wave = 64 / COMPACT_SEG;
    for (uint32_t r0 = (blockIdx.x * WPB + (threadIdx.x >> 6)) * per_wave; r0 < b.n_reads;
         r0 += gridDim.x * WPB * per_wave) {
        const uint32_t r = r0 + seg;
        // the read's count, flags and offsets load together (r clamped to a
        // valid read; lanes past the end copy nothing)
        const uint32_t rr = r < b.n_reads ? r : b.n_reads - 1;
        const uint32_t c0 = b.nam_cnt[rr];
        const uint8_t fl = b.flags[rr];
        const uint64_t ho = b.hit_off[rr], rho = b.rhit_off[rr], no = b.nam_off[rr];
        const uint32_t c = r < b.n_reads ? c0 : 0u;
        const gis_nam *srcn;
        if (fused && !(fl & FLAG_SPILL)) srcn = b.nam_slots + (uint64_t)rr * F_NAMCAP;
        else if (fl & GIS_FLAG_RESCUED) srcn = b.rnam_fin + rho;
        else srcn = b.nam_fin + ho;
        const uint32_t *src = reinterpret_cast<const uint32_t *>(srcn);
        uint32_t *dst = reinterpret_cast<uint32_t *>(b.nams_out + no);
        const uint32_t words = c * (uint32_t)(sizeof(gis_nam) / 4);
        const bool big = words > COMPACT_SEG_WORDS;
        if (!big && words > 0) {
            // every load issues before the first store (indices past the
            // read's words repeat its last word, and store nothing)
            uint32_t v[COMPACT_SEG_WORDS / COMPACT_SEG];
#pragma unroll
            for (uint32_t u = 0; u < COMPACT_SEG_WORDS / COMPACT_SEG; u++) {
                const uint32_t i = j + u * COMPACT_SEG;
                v[u] = src[i < words ? i : words - 1];
            }
#pragma unroll
            for (uint32_t u = 0; u < COMPACT_SEG_WORDS / COMPACT_SEG; u++) {
                const uint32_t i = j + u * COMPACT_SEG;
                if (i < words) dst[i] = v[u];
            }
        }
        uint64_t bigs = __ballot(big && j == 0);
        while (bigs) {
            const uint32_t l = (uint32_t)__builtin_ctzll(bigs);
            bigs &= bigs - 1;
            const uint32_t wl = readlane_u32(words, l);
            const uint32_t *sl = reinterpret_cast<const uint32_t *>(
                (uint64_t)readlane_u32((uint32_t)(uintptr_t)src, l) |
                ((uint64_t)readlane_u32((uint32_t)((uintptr_t)src >> 32), l) << 32));
            uint32_t *dl = reinterpret_cast<uint32_t *>(
                (uint64_t)readlane_u32((uint32_t)(uintptr_t)dst, l) |
                ((uint64_t)readlane_u32((uint32_t)((uintptr_t)dst >> 32), l) << 32));
            // four loads in flight per lane before their stores
            uint32_t i = lane;
            for (; i + 192 < wl; i += 256) {
                const uint32_t v0 = sl[i], v1 = sl[i + 64], v2 = sl[i + 128], v3 = sl[i + 192];
                dl[i] = v0;
                dl[i + 64] = v1;
                dl[i + 128] = v2;
                dl[i + 192] = v3;
            }
            for (; i < wl; i += 64) dl[i] = sl[i];
        }
    }
}

// Ids of the set's reads that a sparse pass visits, in set order within a
// block: SEL_RESCUED (rescued reads) or SEL_DEFERRED (normal reads the
// four-per-wave merge handed on). One atomic per block places a block's ids.
enum { SEL_RESCUED = 0, SEL_DEFERRED = 1, SEL_POOLED = 2 };

__global__ __launch_bounds__(256) void k_select(DevBatch b, ReadSet set, int kind) {
    __shared__ uint32_t wcount[4];
    __shared__ uint32_t block_base;
    const uint32_t lane = lane_id(), wid = threadIdx.x >> 6;
    const uint32_t cnt = set_count(set);
    uint32_t *out = b.sel_ids + (uint64_t)kind * b.n_reads;
    for (uint32_t base = blockIdx.x * 256u; base < cnt; base += gridDim.x * 256u) {
        const uint32_t i = base + threadIdx.x;
        uint32_t r = 0;
        bool m = false;
        if (i < cnt) {
            r = set_at(set, i);
            const bool rescued = (b.flags[r] & GIS_FLAG_RESCUED) != 0;
            m = kind == SEL_RESCUED ? rescued : (!rescued && b.nam_cnt[r] == MERGE_DEFER);
        }
        const uint64_t bal = __ballot(m);
        if (lane == 0) wcount[wid] = (uint32_t)__popcll(bal);
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t t = wcount[0] + wcount[1] + wcount[2] + wcount[3];
            block_base = t ? atomicAdd(&b.ctl->n_sel[kind], t) : 0u;
        }
        __syncthreads();
        uint32_t off = block_base;
        for (uint32_t w = 0; w < wid; w++) off += wcount[w];
        if (m) out[off + mbcnt(bal)] = r;
        __syncthreads();
    }
}

__global__ void k_reset_ctl(DevCtl *ctl) {
    if (threadIdx.x == 0) {
        ctl->n_spill = 0;
        ctl->overflow = 0;
        ctl->q_head[0] = 0;
        ctl->q_head[1] = 0;
        ctl->n_gwork[0] = ctl->n_gwork[1] = 0;
        ctl->n_split[0] = ctl->n_split[1] = 0;
        ctl->n_sel[0] = ctl->n_sel[1] = ctl->n_sel[2] = 0;
        ctl->hit_pool = 0;
        ctl->tot[0] = ctl->tot[1] = ctl->tot[2] = 0;
    }
}

// plain lookup over a list of hashes (the microbenchmark gpu_find,
// gpu_test.cu:186-215, with StrobemerIndex::find semantics)
// find (StrobemerIndex::find) of each hash through its bucket line
__global__ void k_find_flat(DevIndex ix, const uint64_t *hashes, uint64_t n, uint64_t *positions) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Probe p;
    p.key = hashes[i];
    probe_line(ix, p);
    if (probe_small(p)) {
        uint32_t lt, eq;
        probe_small_find(p, lt, eq);
    } else {
        p.pos = p.e - p.s < 0xffffffffull ? dev_find_in(ix, p.key, p.s, p.e, &p.scan) : dev_find(ix, p.key, &p.scan);
    }
    positions[i] = p.pos;
}

// ---------------------------------------------------------------------------
// exclusive scan u32 -> u64 offsets[n+1]

constexpr int SCAN_BLOCK = 256, SCAN_ITEMS = 4, SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

__device__ __forceinline__ uint64_t block_excl_scan_u64(uint64_t v, uint64_t *lds_w, uint64_t *total) {
    const uint32_t lane = lane_id(), wid = threadIdx.x >> 6;
    uint64_t x = v;
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    if (lane == 63) lds_w[wid] = x;
    __syncthreads();
    uint64_t off = 0, tot = 0;
    for (uint32_t w = 0; w < SCAN_BLOCK / 64; w++) {
        if (w < wid) off += lds_w[w];
        tot += lds_w[w];
    }
    __syncthreads();
    *total = tot;
    return off + x - v;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_tiles(const uint32_t *in, uint64_t *out,
                                                           uint64_t n, uint64_t *tile_sums) {
    __shared__ uint64_t lw[SCAN_BLOCK / 64];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint64_t v[SCAN_ITEMS];
    uint64_t s = 0;
    for (int j = 0; j < SCAN_ITEMS; j++) {
        v[j] = base + j < n ? in[base + j] : 0;
        s += v[j];
    }
    uint64_t tot;
    uint64_t ex = block_excl_scan_u64(s, lw, &tot);
    for (int j = 0; j < SCAN_ITEMS; j++) {
        if (base + j < n) out[base + j] = ex;
        ex += v[j];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

// single block: exclusive scan of the tile sums in place; out[n] = total
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_sums(uint64_t *tile_sums, uint64_t ntiles,
                                                          uint64_t *out, uint64_t n) {
    __shared__ uint64_t lw[SCAN_BLOCK / 64];
    uint64_t carry = 0;
    for (uint64_t t0 = 0; t0 < ntiles; t0 += SCAN_BLOCK) {
        const uint64_t t = t0 + threadIdx.x;
        const uint64_t v = t < ntiles ? tile_sums[t] : 0;
        uint64_t tot;
        const uint64_t ex = block_excl_scan_u64(v, lw, &tot);
        if (t < ntiles) tile_sums[t] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) out[n] = carry;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_add(uint64_t *out, uint64_t n,
                                                         const uint64_t *tile_sums) {
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    const uint64_t add = tile_sums[blockIdx.x];
    for (int j = 0; j < SCAN_ITEMS; j++)
        if (base + j < n) out[base + j] += add;
}

// k_scan_sums and k_scan_add in one launch for up to SCAN_FEW_TILES tiles:
// each block adds the sums of the tiles before its own (one wave reads them),
// and the last block writes the total
constexpr unsigned SCAN_FEW_TILES = 256;

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_add_few(uint64_t *out, uint64_t n, const uint64_t *tile_sums,
                                                             uint32_t ntiles) {
    __shared__ uint64_t add_s;
    if (threadIdx.x < 64) {
        uint64_t v = 0;
        for (uint32_t t = threadIdx.x; t < blockIdx.x; t += 64) v += tile_sums[t];
        for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
        if (threadIdx.x == 0) {
            add_s = v;
            if (blockIdx.x == ntiles - 1) out[n] = v + tile_sums[blockIdx.x];
        }
    }
    __syncthreads();
    const uint64_t add = add_s;
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    for (int j = 0; j < SCAN_ITEMS; j++)
        if (base + j < n) out[base + j] += add;
}

// ---------------------------------------------------------------------------
// launch wrappers

static inline unsigned grid_for(uint64_t items, unsigned per_block) {
    return (unsigned)((items + per_block - 1) / per_block);
}

// blocks for a per-read kernel over a set: one wave per read for a full set,
// a bounded grid-stride grid for a spill list (its size is on the device):
// up to 2048 blocks, a whole GPU's worth of waves, since the fused kernel
// can hand on a few percent of a batch (repeat-rich reads with many NAMs),
// and a read's seeding or lookup is a chain of dependent steps
static inline unsigned grid_for_set(const DevBatch &b, const ReadSet &set, unsigned wpb) {
    const unsigned g = grid_for(b.n_reads, wpb);
    if (!set.ids) return g;
    return g < 2048u ? g : 2048u;
}

// the reads a sparse pass visits: the ids k_select wrote
static inline ReadSet selected(const DevBatch &b, int kind) {
    return ReadSet{b.sel_ids + (uint64_t)kind * b.n_reads, &b.ctl->n_sel[kind], b.n_reads};
}

// blocks for a per-read kernel over a selected list (its length is on the
// device): enough waves to run every heavy read of the list at once
static inline unsigned grid_selected(const DevBatch &b) {
    const unsigned g = grid_for(b.n_reads, WPB);
    return g < 2048u ? (g ? g : 1u) : 2048u;
}

static void launch_select(const DevBatch &b, const ReadSet &set, int kind, hipStream_t st) {
    const unsigned g = grid_for(b.n_reads, 256);
    hipLaunchKernelGGL(k_select, dim3(g < 1024u ? (g ? g : 1u) : 1024u), dim3(256), 0, st, b, set, kind);
}

// one thread per bucket: its line (DevIndex::lines) from starts and the table
__global__ __launch_bounds__(256) void k_bucket_lines(const uint4 *rs, const uint64_t *starts, uint64_t nb,
                                                      uint4 *lines) {
    const uint64_t bk = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (bk >= nb) return;
    const uint64_t s = starts[bk], e = starts[bk + 1];
    const uint64_t size = e - s;
    uint4 *line = lines + bk * BUCKET_LINE;
    line[0] = make_uint4((uint32_t)s, (uint32_t)(s >> 32), size < 0xffffffffull ? (uint32_t)size : 0xffffffffu, 0u);
    for (uint32_t j = 1; j < BUCKET_LINE; j++) {
        const uint64_t x = s + j - 1;
        line[j] = x < e ? rs[x] : make_uint4(0u, 0u, 0u, 0u);
    }
}

void launch_bucket_lines(const uint4 *rs, const uint64_t *starts, int bits, uint4 *lines, hipStream_t st) {
    const uint64_t nb = 1ull << bits;
    hipLaunchKernelGGL(k_bucket_lines, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, rs, starts, nb, lines);
}

void launch_reset_ctl(DevCtl *ctl, hipStream_t st) {
    hipLaunchKernelGGL(k_reset_ctl, dim3(1), dim3(64), 0, st, ctl);
}

void launch_seed(const DevBatch &b, const ReadSet &set, const SeedParams &sp, int want_syncmers,
                 int long_reads, hipStream_t st) {
    if (b.n_reads == 0) return;
    const unsigned g = grid_for_set(b, set, SEED_WPB);
    typedef void (*SeedKernel)(DevBatch, ReadSet, SeedParams, int);
    const bool k20 = is_k20_profile(sp);
    const SeedKernel short_k = k20 ? (SeedKernel)k_seed<false, true> : (SeedKernel)k_seed<false, false>;
    const SeedKernel long_k = k20 ? (SeedKernel)k_seed<true, true> : (SeedKernel)k_seed<true, false>;
    hipLaunchKernelGGL(short_k, dim3(g), dim3(64 * SEED_WPB), 0, st, b, set, sp, want_syncmers);
    if (long_reads) hipLaunchKernelGGL(long_k, dim3(g), dim3(64 * SEED_WPB), 0, st, b, set, sp, want_syncmers);
}

void launch_lookup(const DevBatch &b, const ReadSet &set, const DevIndex &ix, int rescue_mode,
                   hipStream_t st) {
    if (b.n_reads == 0) return;
    hipLaunchKernelGGL(k_lookup, dim3(grid_for_set(b, set, WPB)), dim3(64 * WPB), 0, st, b, set, ix,
                       rescue_mode, set.ids ? 1 : 0);
}

void launch_hits(const DevBatch &b, const ReadSet &set, const DevIndex &ix, hipStream_t st) {
    (void)set;
    hipLaunchKernelGGL(k_pool_total, dim3(1), dim3(64), 0, st, b);
    if (b.n_reads == 0) return;
    hipLaunchKernelGGL(k_hits, dim3(grid_selected(b)), dim3(64 * WPB), 0, st, b, selected(b, SEL_POOLED), ix);
}

void launch_rescue_count(const DevBatch &b, const ReadSet &set, const DevIndex &ix,
                         uint32_t rescue_cutoff, uint32_t hard_cutoff, hipStream_t st) {
    if (b.n_reads == 0) return;
    launch_select(b, set, SEL_RESCUED, st);
    hipLaunchKernelGGL(k_rescue_count, dim3(grid_selected(b)), dim3(64 * WPB), 0, st, b,
                       selected(b, SEL_RESCUED), ix, rescue_cutoff, hard_cutoff);
}

// (after launch_rescue_count, which selects the rescued reads)
void launch_rescue_hits(const DevBatch &b, const ReadSet &set, const DevIndex &ix, hipStream_t st) {
    (void)set;
    if (b.n_reads == 0) return;
    hipLaunchKernelGGL(k_rescue_hits, dim3(grid_selected(b)), dim3(64 * WPB), 0, st, b,
                       selected(b, SEL_RESCUED), ix);
}

// normal pass: four reads per wave over the set, then the reads it handed
// on, one per wave; rescue pass: the rescued reads one per wave; then the
// (strand, ref) groups of split reads and their finish
void launch_merge(const DevBatch &b, const ReadSet &set, int rescue, hipStream_t st) {
    if (b.n_reads == 0) return;
    if (!rescue) {
        const unsigned g = grid_for_set(b, set, WPB);
        hipLaunchKernelGGL(k_merge4, dim3((g + RPW - 1) / RPW), dim3(64 * WPB), 0, st, b, set, 0);
        launch_select(b, set, SEL_DEFERRED, st);
        hipLaunchKernelGGL(k_merge, dim3(grid_selected(b)), dim3(64 * WPB), 0, st, b,
                           selected(b, SEL_DEFERRED), 0, 1, 1);
    } else {
        hipLaunchKernelGGL(k_merge, dim3(grid_selected(b)), dim3(64 * WPB), 0, st, b,
                           selected(b, SEL_RESCUED), 1, 0, 1);
    }
    if (b.merge_scratch) {
        hipLaunchKernelGGL(k_merge_groups, dim3(4096), dim3(64), 0, st, b, rescue);
        hipLaunchKernelGGL(k_merge_split_finish, dim3(256), dim3(64 * WPB), 0, st, b, rescue);
    }
}

// fused kernel over queue[0 .. count): FWPB waves per block, LMAX the LDS
// tiling (reads longer than LMAX spill), grid 0 = one wave per entry,
// chunk 0 = static round robin, else dynamic grabs of chunk entries.
template <int LMAX, int FWPB, bool K20>
static void launch_fused_t(const DevBatch &b, const DevIndex &ix, const SeedParams &sp, uint32_t rcut,
                           uint32_t hard, const uint32_t *queue, uint32_t count, unsigned grid,
                           uint32_t chunk, unsigned int *head, hipStream_t st) {
    const unsigned g1 = grid_for(count, FWPB);
    const unsigned g = grid == 0 || grid > g1 ? g1 : grid;
    hipLaunchKernelGGL((k_fused<LMAX, FWPB, K20>), dim3(g), dim3(64 * FWPB), 0, st, b, ix, sp, rcut, hard,
                       queue, count, chunk, head);
}

bool launch_fused(const DevBatch &b, const DevIndex &ix, const SeedParams &sp, uint32_t rcut,
                  uint32_t hard, const uint32_t *queue, uint32_t count, int lmax, int wpb, unsigned grid,
                  uint32_t chunk, unsigned int *head, hipStream_t st) {
    if (count == 0) return true;
#define GIS_FUSED_CASE(LM, W)                                                                      \
    if (lmax == LM && wpb == W) {                                                                  \
        if (is_k20_profile(sp))                                                                    \
            launch_fused_t<LM, W, true>(b, ix, sp, rcut, hard, queue, count, grid, chunk, head, st); \
        else                                                                                       \
            launch_fused_t<LM, W, false>(b, ix, sp, rcut, hard, queue, count, grid, chunk, head, st); \
        return true;                                                                               \
    }
    GIS_FUSED_CASE(160, 2) GIS_FUSED_CASE(160, 4) GIS_FUSED_CASE(160, 8)
    GIS_FUSED_CASE(320, 2) GIS_FUSED_CASE(320, 4) GIS_FUSED_CASE(320, 8)
#undef GIS_FUSED_CASE
    return false;
}

void launch_compact(const DevBatch &b, int fused, hipStream_t st) {
    if (b.n_reads == 0) return;
    hipLaunchKernelGGL(k_compact, dim3(grid_for(b.n_reads, WPB * (64 / COMPACT_SEG))), dim3(64 * WPB), 0, st, b,
                       fused);
}

// The control block to pinned host memory by one wave's stores, so that the
// end of a run takes no copy engine (the engines carry the batches' reads and
// NAMs).
static_assert(sizeof(DevCtl) % 4 == 0 && sizeof(DevCtl) / 4 <= 64, "one word per lane");
__global__ void k_ctl_to_host(const DevCtl *__restrict__ src, DevCtl *__restrict__ dst) {
    const unsigned i = threadIdx.x;
    if (i < sizeof(DevCtl) / 4) reinterpret_cast<unsigned *>(dst)[i] = reinterpret_cast<const unsigned *>(src)[i];
}

void launch_ctl_to_host(const DevCtl *src, DevCtl *dst, hipStream_t st) {
    hipLaunchKernelGGL(k_ctl_to_host, dim3(1), dim3(64), 0, st, src, dst);
}

void launch_find_flat(const DevIndex &ix, const uint64_t *hashes, uint64_t n, uint64_t *positions,
                      hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_find_flat, dim3(grid_for(n, 256)), dim3(256), 0, st, ix, hashes, n, positions);
}

size_t scan_tmp_words(uint64_t n) { return (size_t)grid_for(n ? n : 1, SCAN_TILE) + 1; }

void launch_scan(const uint32_t *in, uint64_t *out, uint64_t n, uint64_t *tmp, hipStream_t st) {
    const unsigned tiles = grid_for(n ? n : 1, SCAN_TILE);
    if (n > 0 && tiles <= SCAN_FEW_TILES) {
        hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(SCAN_BLOCK), 0, st, in, out, n, tmp);
        hipLaunchKernelGGL(k_scan_add_few, dim3(tiles), dim3(SCAN_BLOCK), 0, st, out, n, tmp, (uint32_t)tiles);
        return;
    }
    if (n > 0)
        hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(SCAN_BLOCK), 0, st, in, out, n, tmp);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_BLOCK), 0, st, tmp, (uint64_t)(n ? tiles : 0), out, n);
    if (n > 0) hipLaunchKernelGGL(k_scan_add, dim3(tiles), dim3(SCAN_BLOCK), 0, st, out, n, tmp);
}

}  // namespace gis
