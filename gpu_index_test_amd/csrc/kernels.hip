// gfx950 kernels of the strobealign seeding path.
//
// One wavefront (64 lanes) owns one read in every per-read kernel. Lanes
// work on positions, windows, seeds or hits of that read in parallel; the
// wave synchronises through LDS (seeding) or through its own global scratch
// regions (hits, merge). Reads are independent, so there is no inter-wave
// communication at all.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"

namespace gis {

// ---------------------------------------------------------------------------
// wave helpers

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// wave-local ordering of global-memory scratch written by other lanes of the
// same wave (all lanes of a workgroup share one CU's L1)
__device__ __forceinline__ void wave_sync_mem() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
    return x;
}

__device__ __forceinline__ uint32_t wave_max(uint32_t x) {
    for (int d = 32; d >= 1; d >>= 1) {
        uint32_t y = __shfl_xor(x, d, 64);
        x = x > y ? x : y;
    }
    return x;
}

__device__ __forceinline__ uint32_t wave_min(uint32_t x) {
    for (int d = 32; d >= 1; d >>= 1) {
        uint32_t y = __shfl_xor(x, d, 64);
        x = x < y ? x : y;
    }
    return x;
}

__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v) {
    const uint32_t lane = lane_id();
    uint32_t x = v;
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    return x - v;
}

// ---------------------------------------------------------------------------
// seeding: canonical syncmers and query randstrobes of one read

struct SeedBufs {
    uint8_t *codes;     // 2-bit base codes, 4 = not ACGT
    uint8_t *st;        // per k-mer window: anchor state / ST_UNKNOWN / ST_INVALID
    uint64_t *sh;       // canonical s-mer hash per s-mer start
    uint64_t *sy_hash;  // syncmer list
    uint32_t *sy_pos;
    uint64_t *nmask;    // bit i set when base i is not ACGT
};

// any non-ACGT base in [a, b] (b - a < 64)
__device__ __forceinline__ bool any_n(const uint64_t *nm, int a, int b) {
    const int wa = a >> 6, wb = b >> 6;
    if (wa == wb) {
        const uint64_t mask = (~0ULL >> (63 - (b - a))) << (a & 63);
        return (nm[wa] & mask) != 0;
    }
    const uint64_t m1 = nm[wa] >> (a & 63);
    const uint64_t m2 = nm[wb] & (~0ULL >> (63 - (b & 63)));
    return (m1 | m2) != 0;
}

// The whole wave calls this for the same read. Outputs: optional syncmer
// list (out_sy_*), randstrobes (forward then reverse complement, as
// randstrobes_query orders them, step2/randstrobes.cu:195-241).
__device__ __forceinline__ void seed_one(const char *__restrict__ seq, int L, const SeedParams &P,
                                         const SeedBufs &B, uint64_t *out_sy_hash,
                                         uint64_t *out_sy_pos, uint64_t *rs_hash,
                                         uint32_t *rs_start, uint32_t *rs_end,
                                         uint32_t *n_sync_out, uint32_t *n_rs_out) {
    const int lane = (int)lane_id();
    const int k = P.k, s = P.s, w = P.w;

    // 1. base codes and the N mask (step2/randstrobes.cu:21-22, 70-76)
    for (int b0 = 0; b0 < L; b0 += 64) {
        const int i = b0 + lane;
        int c = 4;
        if (i < L) {
            c = nt4((unsigned char)seq[i]);
            B.codes[i] = (uint8_t)c;
        }
        const uint64_t m = __ballot(i < L && c > 3);
        if (lane == 0) B.nmask[b0 >> 6] = m;
    }
    wave_sync();

    // 2. canonical s-mer hashes, one s-mer per lane (step2/randstrobes.cu:23-32)
    const int ns = L - s + 1;
    for (int j0 = 0; j0 < ns; j0 += 64) {
        const int j = j0 + lane;
        if (j < ns) {
            uint64_t h = 0;
            if (!any_n(B.nmask, j, j + s - 1)) {
                uint64_t f = 0, r = 0;
                for (int m = 0; m < s; m++) {
                    const uint64_t c = B.codes[j + m];
                    f = (f << 2) | c;
                    r |= (3 - c) << (2 * m);
                }
                h = xxh64_u64(f < r ? f : r);
            }
            B.sh[j] = h;
        }
    }
    wave_sync();

    // 3a. window states that need no history
    const int nk = L - k + 1;
    for (int p0 = 0; p0 < nk; p0 += 64) {
        const int p = p0 + lane;
        if (p < nk) {
            uint8_t st;
            if (any_n(B.nmask, p, p + k - 1)) {
                st = ST_INVALID;
            } else {
                const bool seg_first = (p == 0) || B.codes[p - 1] > 3;
                st = anchor_state(B.sh, p, w, seg_first);
            }
            B.st[p] = st;
        }
    }
    wave_sync();

    // 3b. resolve tied windows from the nearest anchor and compact syncmers;
    // a window is a syncmer when its minimum sits at offset t-1
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
            uint64_t f = 0, r = 0;
            for (int m = 0; m < k; m++) {
                const uint64_t c = B.codes[p + m];
                f = (f << 2) | c;
                r |= (3 - c) << (2 * m);
            }
            f &= P.kmask;
            const uint64_t hk = xxh64_u64(f < r ? f : r);
            B.sy_hash[idx] = hk;
            B.sy_pos[idx] = (uint32_t)p;
            if (out_sy_hash) {
                out_sy_hash[idx] = hk;
                out_sy_pos[idx] = (uint64_t)p;
            }
        }
        nsync += (uint32_t)__popcll(bal);
    }
    wave_sync();

    // 4. randstrobes (RandstrobeIterator::get, step2/randstrobes.cu:136-159);
    // the reverse-complement pass walks the forward syncmers backwards with
    // mirrored positions L - pos - k (step2/randstrobes.cu:221-239)
    const uint32_t n = nsync;
    const uint32_t n_fwd = ((uint32_t)L >= P.w_max && n > P.w_min) ? n - P.w_min : 0;
    for (uint32_t i0 = 0; i0 < n_fwd; i0 += 64) {
        const uint32_t i = i0 + (uint32_t)lane;
        if (i < n_fwd) {
            const uint32_t wend = (i + P.w_max < n - 1) ? i + P.w_max : n - 1;
            {
                const uint64_t h1 = B.sy_hash[i];
                const uint32_t p1 = B.sy_pos[i];
                const uint64_t maxp = (uint64_t)p1 + P.max_dist_u;
                uint32_t best = i;
                uint64_t minv = UINT64_MAX;
                for (uint32_t j = i + P.w_min; j <= wend && (uint64_t)B.sy_pos[j] <= maxp; j++) {
                    const uint64_t c = (uint64_t)__popcll((h1 ^ B.sy_hash[j]) & P.q);
                    if (c < minv) { minv = c; best = j; }
                }
                rs_hash[i] = h1 + B.sy_hash[best];
                rs_start[i] = p1;
                rs_end[i] = B.sy_pos[best] + (uint32_t)k;
            }
            {
                const uint32_t a = n - 1 - i;
                const uint64_t h1 = B.sy_hash[a];
                const uint32_t p1 = (uint32_t)L - B.sy_pos[a] - (uint32_t)k;
                const uint64_t maxp = (uint64_t)p1 + P.max_dist_u;
                uint32_t best = a;
                uint64_t minv = UINT64_MAX;
                for (uint32_t jj = i + P.w_min; jj <= wend; jj++) {
                    const uint32_t bidx = n - 1 - jj;
                    const uint32_t pb = (uint32_t)L - B.sy_pos[bidx] - (uint32_t)k;
                    if ((uint64_t)pb > maxp) break;
                    const uint64_t c = (uint64_t)__popcll((h1 ^ B.sy_hash[bidx]) & P.q);
                    if (c < minv) { minv = c; best = bidx; }
                }
                rs_hash[n_fwd + i] = h1 + B.sy_hash[best];
                rs_start[n_fwd + i] = p1;
                rs_end[n_fwd + i] = (uint32_t)L - B.sy_pos[best];
            }
        }
    }
    *n_sync_out = n;
    *n_rs_out = 2 * n_fwd;
}

struct alignas(8) SeedLds {
    uint64_t sh[LMAX_LDS];
    uint64_t sy_hash[LMAX_LDS];
    uint64_t nmask[LMAX_LDS / 64 + 2];
    uint32_t sy_pos[LMAX_LDS];
    uint8_t codes[LMAX_LDS];
    uint8_t st[LMAX_LDS];
};

__device__ __forceinline__ void store_seed_out(const DevBatch &b, uint32_t r, int want_sy,
                                               uint32_t nsync, uint32_t nrs) {
    if (lane_id() == 0) {
        b.rs_cnt[r] = nrs;
        b.rs_nfwd[r] = nrs / 2;
        if (want_sy) b.sy_cnt[r] = nsync;
    }
}

__global__ __launch_bounds__(64 * SEED_WPB) void k_seed(DevBatch b, SeedParams P, int want_sy) {
    __shared__ SeedLds lds[SEED_WPB];
    const int wid = (int)(threadIdx.x >> 6);
    const uint32_t r = blockIdx.x * SEED_WPB + (uint32_t)wid;
    if (r >= b.n_reads) return;
    const int L = (int)b.read_len[r];
    if (L > LMAX_LDS) return;  // long reads go through k_seed_long
    SeedLds &S = lds[wid];
    SeedBufs B{S.codes, S.st, S.sh, S.sy_hash, S.sy_pos, S.nmask};
    uint64_t *osh = want_sy ? b.sy_hash + b.sy_off[r] : nullptr;
    uint64_t *osp = want_sy ? b.sy_pos + b.sy_off[r] : nullptr;
    const uint64_t ro = b.rs_off[r];
    uint32_t nsync = 0, nrs = 0;
    seed_one(b.seqs + b.read_off[r], L, P, B, osh, osp, b.rs_hash + ro, b.rs_start + ro,
             b.rs_end + ro, &nsync, &nrs);
    store_seed_out(b, r, want_sy, nsync, nrs);
}

// long reads: the same wave algorithm with its buffers in global scratch
__host__ __device__ inline uint64_t align8(uint64_t x) { return (x + 7) & ~7ULL; }

__global__ __launch_bounds__(64 * SEED_WPB) void k_seed_long(DevBatch b, SeedParams P, int want_sy) {
    const int wid = (int)(threadIdx.x >> 6);
    const uint32_t r = blockIdx.x * SEED_WPB + (uint32_t)wid;
    if (r >= b.n_reads) return;
    const int L = (int)b.read_len[r];
    if (L <= LMAX_LDS) return;
    uint8_t *base = b.long_scratch + b.long_off[r];
    const uint64_t Lu = (uint64_t)L;
    SeedBufs B;
    B.sh = (uint64_t *)base;
    B.sy_hash = (uint64_t *)(base + 8 * Lu);
    B.nmask = (uint64_t *)(base + 16 * Lu);
    B.sy_pos = (uint32_t *)(base + 16 * Lu + 8 * (Lu / 64 + 2));
    B.codes = base + 16 * Lu + 8 * (Lu / 64 + 2) + align8(4 * Lu);
    B.st = B.codes + align8(Lu);
    uint64_t *osh = want_sy ? b.sy_hash + b.sy_off[r] : nullptr;
    uint64_t *osp = want_sy ? b.sy_pos + b.sy_off[r] : nullptr;
    const uint64_t ro = b.rs_off[r];
    uint32_t nsync = 0, nrs = 0;
    seed_one(b.seqs + b.read_off[r], L, P, B, osh, osp, b.rs_hash + ro, b.rs_start + ro,
             b.rs_end + ro, &nsync, &nrs);
    store_seed_out(b, r, want_sy, nsync, nrs);
}

// per-read capacities: syncmers <= L-k+1, randstrobes <= 2 (L-k+1-w_min)
__global__ void k_capacity(const uint32_t *lens, uint32_t n, SeedParams P, uint32_t *cap_rs,
                           uint32_t *cap_sy) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int L = (int)lens[r];
    const int nk = L - P.k + 1;
    const int nf = nk - (int)P.w_min;
    cap_sy[r] = nk > 0 ? (uint32_t)nk : 0u;
    cap_rs[r] = (nf > 0 && (uint32_t)L >= P.w_max) ? 2u * (uint32_t)nf : 0u;
}

// ---------------------------------------------------------------------------
// index access (StrobemerIndex, step2/index.hpp:57-147)

__device__ __forceinline__ uint64_t rec_hash(const DevIndex &ix, uint64_t x) {
    return reinterpret_cast<const uint64_t *>(ix.rs)[2 * x];
}

// first index of the bucket range holding key, or GIS_NOT_FOUND
// (StrobemerIndex::find, step2/index.hpp:57-81: linear scan of a small bucket,
// lower_bound otherwise; both return the first equal entry). *probes counts
// the table records read.
__device__ __forceinline__ uint64_t dev_find(const DevIndex &ix, uint64_t key, uint32_t *probes) {
    const uint32_t bk = (uint32_t)(key >> (64 - ix.bits));
    const uint64_t s = ix.starts[bk], e = ix.starts[bk + 1];
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
        if (lo < e) {
            pr++;
            if (rec_hash(ix, lo) == key) res = lo;
        }
    }
    *probes = pr;
    return res;
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
// kept hits, and writes them when out != nullptr.
__device__ __forceinline__ uint32_t walk_run(const DevIndex &ix, uint64_t pos, uint64_t key,
                                             int qs, int qe, uint32_t strand_bit, DevHit *out,
                                             uint32_t *records = nullptr) {
    const int qspan = qe - qs;
    int min_diff = 0x7fffffff;
    uint32_t c = 0, rec = 0;
    for (uint64_t x = pos; x < ix.n; x++) {
        const uint4 v = ix.rs[x];
        rec++;
        const uint64_t h = (uint64_t)v.x | ((uint64_t)v.y << 32);
        if (h != key) break;
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
    if (records) *records = rec;
    return c;
}

// ---------------------------------------------------------------------------
// lookup: find + filter + hit count per randstrobe; nonrepetitive fraction
// and the rescue decision per read (find_nams, step2/nam.cu:553-732, and the
// caller's test, step2/main.cu:3209-3216)

__global__ __launch_bounds__(64 * WPB) void k_lookup(DevBatch b, DevIndex ix, int rescue_mode) {
    const uint32_t r = blockIdx.x * WPB + (threadIdx.x >> 6);
    if (r >= b.n_reads) return;
    const uint32_t lane = lane_id();
    const uint32_t n = b.rs_cnt[r];
    const uint64_t base = b.rs_off[r];
    const uint32_t half = b.rs_nfwd[r];
    uint32_t found = 0, good = 0, hits = 0, bytes = 0;
    for (uint32_t i0 = 0; i0 < n; i0 += 64) {
        const uint32_t i = i0 + lane;
        uint32_t f = 0, g = 0, hc = 0, by = 0;
        if (i < n) {
            const uint64_t key = b.rs_hash[base + i];
            uint32_t probes = 0, records = 0;
            const uint64_t pos = dev_find(ix, key, &probes);
            // query in (16) + result out (12) + bucket bounds (16) + search records
            by = 44 + 16 * probes;
            if (pos != GIS_NOT_FOUND) {
                f = 1;
                by += pos + ix.filter_cutoff < ix.n ? 8 : 0;
                if (!dev_filtered(ix, pos, key)) {
                    g = 1;
                    hc = walk_run(ix, pos, key, (int)b.rs_start[base + i], (int)b.rs_end[base + i],
                                  i >= half ? 0x80000000u : 0u, nullptr, &records);
                    by += 16 * records;
                }
            }
            b.pos[base + i] = pos;
            b.hcnt[base + i] = hc;
        }
        found += wave_sum(f);
        good += wave_sum(g);
        hits += wave_sum(hc);
        bytes += wave_sum(by);
    }
    if (lane == 0) {
        b.lk_bytes[r] = bytes;
        const float frac = found > 0 ? (float)good / (float)found : 1.0f;
        const bool rescue = rescue_mode == 2 || (rescue_mode == 0 && ((double)frac < 0.7 || hits == 0));
        b.frac[r] = frac;
        b.flags[r] = (uint8_t)((rescue ? GIS_FLAG_RESCUED : 0) | (n == 0 ? GIS_FLAG_NO_RANDSTROBES : 0));
        b.hit_total[r] = rescue ? 0u : hits;  // rescued reads discard these hits
    }
}

// write the hits of non-rescued reads at hit_off (same walk as k_lookup)
__global__ __launch_bounds__(64 * WPB) void k_hits(DevBatch b, DevIndex ix) {
    const uint32_t r = blockIdx.x * WPB + (threadIdx.x >> 6);
    if (r >= b.n_reads) return;
    if (b.flags[r] & GIS_FLAG_RESCUED) return;
    const uint32_t lane = lane_id();
    const uint32_t n = b.rs_cnt[r];
    const uint64_t base = b.rs_off[r];
    const uint32_t half = b.rs_nfwd[r];
    DevHit *out = b.hits + b.hit_off[r];
    uint32_t carry = 0;
    for (uint32_t i0 = 0; i0 < n; i0 += 64) {
        const uint32_t i = i0 + lane;
        const uint32_t hc = i < n ? b.hcnt[base + i] : 0u;
        const uint32_t ex = wave_excl_scan(hc);
        if (hc > 0) {
            const uint64_t key = b.rs_hash[base + i];
            walk_run(ix, b.pos[base + i], key, (int)b.rs_start[base + i], (int)b.rs_end[base + i],
                     i >= half ? 0x80000000u : 0u, out + carry + ex);
        }
        carry += wave_sum(hc);
    }
}


// ---------------------------------------------------------------------------
// rescue (find_nams_rescue, step2/nam.cu:927-984)

// Per rescued read: the run length (get_count) of every found seed goes to
// hcnt. A seed is taken when count <= hard_cutoff and (count <= rescue_cutoff
// or fewer than 5 seeds of its strand precede it in (count, query_start,
// query_end) order) -- exactly the prefix the reference's sorted loop keeps
// (step2/nam.cu:957-972). Taken seeds later walk their runs in query_start
// order, which is the seeds' own order within a strand. rcnt gets the number
// of hits each seed adds.
__global__ __launch_bounds__(64 * WPB) void k_rescue_count(DevBatch b, DevIndex ix,
                                                         uint32_t rescue_cutoff,
                                                         uint32_t hard_cutoff) {
    const uint32_t r = blockIdx.x * WPB + (threadIdx.x >> 6);
    if (r >= b.n_reads) return;
    const uint32_t lane = lane_id();
    if (!(b.flags[r] & GIS_FLAG_RESCUED)) {
        if (lane == 0) b.rhit_total[r] = 0;
        return;
    }
    const uint32_t n = b.rs_cnt[r];
    const uint64_t base = b.rs_off[r];
    const uint32_t half = b.rs_nfwd[r];
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
        if (i < n) {
            const uint32_t c = b.hcnt[base + i];
            bool taken = false;
            if (c > 0 && c <= hard_cutoff) {
                taken = c <= rescue_cutoff;
                if (!taken) {
                    const uint32_t lo = i < half ? 0 : half, hi = i < half ? half : n;
                    const uint32_t qs = b.rs_start[base + i], qe = b.rs_end[base + i];
                    uint32_t rank = 0;
                    for (uint32_t j = lo; j < hi && rank < 5; j++) {
                        const uint32_t cj = b.hcnt[base + j];
                        if (cj == 0) continue;
                        const uint32_t qsj = b.rs_start[base + j], qej = b.rs_end[base + j];
                        const bool less = cj < c || (cj == c && (qsj < qs || (qsj == qs && qej < qe)));
                        rank += less ? 1u : 0u;
                    }
                    taken = rank < 5;
                }
            }
            if (taken)
                rh = walk_run(ix, b.pos[base + i], b.rs_hash[base + i], (int)b.rs_start[base + i],
                              (int)b.rs_end[base + i], 0u, nullptr);
            b.rcnt[base + i] = rh;
        }
        total += wave_sum(rh);
    }
    if (lane == 0) b.rhit_total[r] = total;
}

__global__ __launch_bounds__(64 * WPB) void k_rescue_hits(DevBatch b, DevIndex ix) {
    const uint32_t r = blockIdx.x * WPB + (threadIdx.x >> 6);
    if (r >= b.n_reads) return;
    if (!(b.flags[r] & GIS_FLAG_RESCUED)) return;
    const uint32_t lane = lane_id();
    const uint32_t n = b.rs_cnt[r];
    const uint64_t base = b.rs_off[r];
    const uint32_t half = b.rs_nfwd[r];
    DevHit *out = b.rhits + b.rhit_off[r];
    uint32_t carry = 0;
    for (uint32_t i0 = 0; i0 < n; i0 += 64) {
        const uint32_t i = i0 + lane;
        const uint32_t rh = i < n ? b.rcnt[base + i] : 0u;
        const uint32_t ex = wave_excl_scan(rh);
        if (rh > 0)
            walk_run(ix, b.pos[base + i], b.rs_hash[base + i], (int)b.rs_start[base + i],
                     (int)b.rs_end[base + i], i >= half ? 0x80000000u : 0u, out + carry + ex);
        carry += wave_sum(rh);
    }
}

// ---------------------------------------------------------------------------
// NAM merge. Hits of a read arrive strand by strand, seeds in query_start
// order and each run in table order; within one (strand, ref) group that is
// (query_start, ref_start) order, the order merge_hits_into_nams sorts into
// (step2/nam.cu:373). Hits are grouped by (strand, ref_id) with a stable
// sort, then each group is merged by one lane: every hit extends or is
// contained in the first open NAM (in creation order) that accepts it, or
// opens a new NAM (step2/nam.cu:391-447). NAMs whose query_end lies before
// the current query_start can never accept a hit again and are dropped from
// the open list; when they are written out does not change any field.
// Creation order inside a group is (query_start, ref_start) order, so
// writing groups in key order gives the canonical NAM order directly.

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

__device__ void bitonic_global(uint64_t *keys, uint32_t P) {
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

__global__ __launch_bounds__(64 * WPB) void k_merge(DevBatch b, int rescue, int k) {
    const uint32_t r = blockIdx.x * WPB + (threadIdx.x >> 6);
    if (r >= b.n_reads) return;
    const bool rescued = (b.flags[r] & GIS_FLAG_RESCUED) != 0;
    if (rescued != (rescue != 0)) return;
    const uint32_t lane = lane_id();
    const uint64_t hb = rescue ? b.rhit_off[r] : b.hit_off[r];
    const uint32_t nh = (uint32_t)((rescue ? b.rhit_off[r + 1] : b.hit_off[r + 1]) - hb);
    const DevHit *hits = (rescue ? b.rhits : b.hits) + hb;
    uint64_t *keys = (rescue ? b.rkeys : b.keys) + 2 * hb;
    uint32_t *aux = (rescue ? b.raux : b.aux) + 2 * hb + r;
    gis_nam *ws = (rescue ? b.rnam_ws : b.nam_ws) + hb;
    gis_nam *fin = (rescue ? b.rnam_fin : b.nam_fin) + hb;
    if (nh == 0) {
        if (lane == 0) b.nam_cnt[r] = 0;
        return;
    }
    // keys = (is_rc << 31 | ref_id) << 32 | hit index; sort only when the
    // hits are not already grouped in key order
    uint32_t unsorted = 0;
    for (uint32_t i0 = 0; i0 < nh; i0 += 64) {
        const uint32_t i = i0 + lane;
        if (i < nh) {
            const uint32_t rsd = hits[i].ref_strand;
            keys[i] = ((uint64_t)rsd << 32) | i;
            if (i > 0 && hits[i - 1].ref_strand > rsd) unsorted = 1;
        }
    }
    unsorted = wave_max(unsorted);
    if (unsorted) {
        if (nh <= 64) {
            wave_sync_mem();
            uint64_t key = lane < nh ? keys[lane] : UINT64_MAX;
            bitonic_regs(key);
            if (lane < nh) keys[lane] = key;
        } else {
            uint32_t P = 1;
            while (P < nh) P <<= 1;
            for (uint32_t i = nh + lane; i < P; i += 64) keys[i] = UINT64_MAX;
            wave_sync_mem();
            bitonic_global(keys, P);
        }
    }
    wave_sync_mem();
    // group starts -> aux[0..G]; open lists live at aux[nh + 1 + group start]
    uint32_t G = 0;
    for (uint32_t i0 = 0; i0 < nh; i0 += 64) {
        const uint32_t i = i0 + lane;
        bool head = false;
        if (i < nh) head = (i == 0) || ((keys[i] >> 32) != (keys[i - 1] >> 32));
        const uint64_t bal = __ballot(head);
        if (head) aux[G + mbcnt(bal)] = i;
        G += (uint32_t)__popcll(bal);
    }
    if (lane == 0) aux[G] = nh;
    wave_sync_mem();
    uint32_t *open_base = aux + nh + 1;
    for (uint32_t g0 = 0; g0 < G; g0 += 64) {
        const uint32_t g = g0 + lane;
        if (g < G) {
            const uint32_t gs = aux[g], ge = aux[g + 1];
            uint32_t *open = open_base + gs;
            gis_nam *gw = ws + gs;
            uint32_t n_open = 0, c = 0;
            int cur_qs = -1;
            for (uint32_t x = gs; x < ge; x++) {
                const DevHit h = hits[(uint32_t)keys[x]];
                const int qs = h.query_start, qe = h.query_end, rs = h.ref_start, re = h.ref_end;
                if (qs != cur_qs) {
                    uint32_t w2 = 0;
                    for (uint32_t oi = 0; oi < n_open; oi++) {
                        const uint32_t o = open[oi];
                        if (gw[o].query_end >= qs) open[w2++] = o;
                    }
                    n_open = w2;
                    cur_qs = qs;
                }
                bool added = false;
                for (uint32_t oi = 0; oi < n_open; oi++) {
                    gis_nam &o = gw[open[oi]];
                    if (o.query_prev_hit_startpos < qs && qs <= o.query_end &&
                        o.ref_prev_hit_startpos < rs && rs <= o.ref_end) {
                        if (qe > o.query_end && re > o.ref_end) {
                            o.query_end = qe;
                            o.ref_end = re;
                            o.query_prev_hit_startpos = qs;
                            o.ref_prev_hit_startpos = rs;
                            o.n_hits++;
                            added = true;
                            break;
                        } else if (qe <= o.query_end && re <= o.ref_end) {
                            o.query_prev_hit_startpos = qs;
                            o.ref_prev_hit_startpos = rs;
                            o.n_hits++;
                            added = true;
                            break;
                        }
                    }
                }
                if (!added) {
                    gis_nam &o = gw[c];
                    o.query_start = qs;
                    o.query_end = qe;
                    o.query_prev_hit_startpos = qs;
                    o.ref_start = rs;
                    o.ref_end = re;
                    o.ref_prev_hit_startpos = rs;
                    o.n_hits = 1;
                    open[n_open++] = c;
                    c++;
                }
            }
            open[0] = c;  // the group's NAM count (open list no longer needed)
        }
    }
    wave_sync_mem();
    // ranks of groups -> final NAMs with nam_id = rank, score filled in
    uint32_t carry = 0;
    for (uint32_t g0 = 0; g0 < G; g0 += 64) {
        const uint32_t g = g0 + lane;
        uint32_t c = 0, gs = 0;
        if (g < G) {
            gs = aux[g];
            c = open_base[gs];
        }
        const uint32_t rank = carry + wave_excl_scan(c);
        if (g < G) {
            const uint32_t rsd = (uint32_t)(keys[gs] >> 32);
            for (uint32_t m = 0; m < c; m++) {
                gis_nam o = ws[gs + m];
                o.nam_id = (int32_t)(rank + m);
                o.ref_id = (int32_t)(rsd & 0x7fffffffu);
                o.is_rc = (uint8_t)(rsd >> 31);
                o.pad_[0] = o.pad_[1] = o.pad_[2] = 0;
                o.score = nam_score(o.query_start, o.query_end, o.ref_start, o.ref_end, o.n_hits);
                fin[rank + m] = o;
            }
        }
        carry += wave_sum(c);
    }
    if (lane == 0) b.nam_cnt[r] = carry;
}

// dense output: NAMs of read r to nams_out[nam_off[r] ..]
__global__ __launch_bounds__(64 * WPB) void k_compact(DevBatch b) {
    const uint32_t r = blockIdx.x * WPB + (threadIdx.x >> 6);
    if (r >= b.n_reads) return;
    const uint32_t lane = lane_id();
    const bool rescued = (b.flags[r] & GIS_FLAG_RESCUED) != 0;
    const uint32_t c = b.nam_cnt[r];
    const uint32_t *src = reinterpret_cast<const uint32_t *>(
        rescued ? b.rnam_fin + b.rhit_off[r] : b.nam_fin + b.hit_off[r]);
    uint32_t *dst = reinterpret_cast<uint32_t *>(b.nams_out + b.nam_off[r]);
    const uint32_t words = c * (uint32_t)(sizeof(gis_nam) / 4);
    for (uint32_t i = lane; i < words; i += 64) dst[i] = src[i];
}

// plain lookup over a list of hashes (the microbenchmark gpu_find,
// gpu_test.cu:186-215, with StrobemerIndex::find semantics)
__global__ void k_find_flat(DevIndex ix, const uint64_t *hashes, uint64_t n, uint64_t *positions) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    positions[i] = dev_find(ix, hashes[i]);
}

// ---------------------------------------------------------------------------
// exclusive scan u32 -> u64 offsets[n+1]

constexpr int SCAN_BLOCK = 256, SCAN_ITEMS = 4, SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

__device__ __forceinline__ uint64_t block_excl_scan_u64(uint64_t v, uint64_t *lds_w, uint64_t *total) {
    // v per thread; 4 waves
    const uint32_t lane = lane_id(), wid = threadIdx.x >> 6;
    uint64_t x = v;
    for (int d = 1; d < 64; d <<= 1) {
        uint64_t y = __shfl_up(x, d, 64);
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

// ---------------------------------------------------------------------------
// launch wrappers

static inline unsigned grid_for(uint64_t items, unsigned per_block) {
    return (unsigned)((items + per_block - 1) / per_block);
}

void launch_seed(const DevBatch &b, const SeedParams &sp, int want_syncmers, int long_reads,
                 hipStream_t st) {
    if (b.n_reads == 0) return;
    hipLaunchKernelGGL(k_seed, dim3(grid_for(b.n_reads, SEED_WPB)), dim3(64 * SEED_WPB), 0, st, b, sp,
                       want_syncmers);
    if (long_reads)
        hipLaunchKernelGGL(k_seed_long, dim3(grid_for(b.n_reads, SEED_WPB)), dim3(64 * SEED_WPB), 0,
                           st, b, sp, want_syncmers);
}

void launch_lookup(const DevBatch &b, const DevIndex &ix, int rescue_mode, hipStream_t st) {
    if (b.n_reads == 0) return;
    hipLaunchKernelGGL(k_lookup, dim3(grid_for(b.n_reads, WPB)), dim3(64 * WPB), 0, st, b, ix,
                       rescue_mode);
}

void launch_hits(const DevBatch &b, const DevIndex &ix, hipStream_t st) {
    if (b.n_reads == 0) return;
    hipLaunchKernelGGL(k_hits, dim3(grid_for(b.n_reads, WPB)), dim3(64 * WPB), 0, st, b, ix);
}

void launch_merge(const DevBatch &b, int rescue, int k, hipStream_t st) {
    if (b.n_reads == 0) return;
    hipLaunchKernelGGL(k_merge, dim3(grid_for(b.n_reads, WPB)), dim3(64 * WPB), 0, st, b, rescue, k);
}

void launch_rescue_count(const DevBatch &b, const DevIndex &ix, uint32_t rescue_cutoff,
                         uint32_t hard_cutoff, hipStream_t st) {
    if (b.n_reads == 0) return;
    hipLaunchKernelGGL(k_rescue_count, dim3(grid_for(b.n_reads, WPB)), dim3(64 * WPB), 0, st, b, ix,
                       rescue_cutoff, hard_cutoff);
}

void launch_rescue_hits(const DevBatch &b, const DevIndex &ix, uint32_t, uint32_t, hipStream_t st) {
    if (b.n_reads == 0) return;
    hipLaunchKernelGGL(k_rescue_hits, dim3(grid_for(b.n_reads, WPB)), dim3(64 * WPB), 0, st, b, ix);
}

void launch_compact(const DevBatch &b, hipStream_t st) {
    if (b.n_reads == 0) return;
    hipLaunchKernelGGL(k_compact, dim3(grid_for(b.n_reads, WPB)), dim3(64 * WPB), 0, st, b);
}

void launch_find_flat(const DevIndex &ix, const uint64_t *hashes, uint64_t n, uint64_t *positions,
                      hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_find_flat, dim3(grid_for(n, 256)), dim3(256), 0, st, ix, hashes, n, positions);
}

void launch_rs_capacity(const uint32_t *lens, uint32_t n, const SeedParams &sp, uint32_t *cap_rs,
                        uint32_t *cap_sy, hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_capacity, dim3(grid_for(n, 256)), dim3(256), 0, st, lens, n, sp, cap_rs, cap_sy);
}

size_t scan_tmp_words(uint64_t n) { return (size_t)grid_for(n ? n : 1, SCAN_TILE) + 1; }

void launch_scan(const uint32_t *in, uint64_t *out, uint64_t n, uint64_t *tmp, hipStream_t st) {
    const unsigned tiles = grid_for(n ? n : 1, SCAN_TILE);
    if (n > 0)
        hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(SCAN_BLOCK), 0, st, in, out, n, tmp);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_BLOCK), 0, st, tmp, (uint64_t)(n ? tiles : 0), out, n);
    if (n > 0) hipLaunchKernelGGL(k_scan_add, dim3(tiles), dim3(SCAN_BLOCK), 0, st, out, n, tmp);
}

}  // namespace gis
