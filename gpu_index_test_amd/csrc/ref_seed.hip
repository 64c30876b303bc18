// Device seeding of the reference for the index build: the RefRandstrobe
// records of every contig, as the reference's RandstrobeGenerator makes
// them over each contig's syncmers (step2/randstrobes.cu:18-79,136-190;
// count/assign_randstrobes at step2/index.cu:246-311).
//
// The contigs sit in HBM back to back with one 'N' after each, so a contig
// boundary acts as an N run: no window spans two contigs and the first
// window after a boundary starts a segment. The scan runs as passes over
// the whole genome, each a grid over positions:
//   1. bases packed 2 bits each, with a mask of non-ACGT bases (ballots);
//   2. the canonical s-mer hash at every position;
//   3. the state of every k-mer window that needs no history (segment
//      starts and windows with a unique minimum below ~0, gis_core.h);
//   4. each maximal run of windows with tied minima replayed from the known
//      window before it by one lane (slide, gis_core.h) -- tied runs are
//      rare and short except in tandem repeats, and a run's replay is
//      sequential in the reference too;
//   5. syncmer positions compacted per 4096-window tile (scan of tile
//      counts), then the canonical k-mer hash of each syncmer;
//   6. per contig, the range of its syncmers; then one lane per syncmer
//      picks its partner among the next w_min..w_max syncmers of the contig
//      within max_dist (the smallest popcount((h_i ^ h_j) & q), earliest on
//      ties) and writes its record. A contig's last w_min syncmers start no
//      randstrobe; a contig shorter than w_max gets zero records in its
//      slots, as the reference counts them but never assigns them.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "index_build.h"
#include "kernels.h"
#include "seed_dev.h"
#include "wave.h"

namespace gis {

namespace {

constexpr int TILE_THREADS = 256, TILE_WIN = 16, TILE = TILE_THREADS * TILE_WIN;

struct Dev {
    void *p = nullptr;
    ~Dev() {
        if (p) (void)hipFree(p);
    }
};

template <typename T>
T *dalloc(Dev &d, uint64_t n) {
    const hipError_t e = hipMalloc(&d.p, std::max<uint64_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) {
        d.p = nullptr;
        throw BuildError{e, "hipMalloc (reference seeding)"};
    }
    return static_cast<T *>(d.p);
}

void check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw BuildError{e, what};
}

unsigned grid(uint64_t items, unsigned per_block) {
    const uint64_t g = (items + per_block - 1) / per_block;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(g, 1u << 20));
}

#define GRID_STRIDE(i, n) \
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * blockDim.x)

// 1. one wave per 64 bases: pk[2w], pk[2w+1] and nmask[w]
__global__ void k_ref_pack(const char *__restrict__ seq, uint64_t T, uint64_t nwords, uint64_t *__restrict__ pk,
                           uint64_t *__restrict__ nmask) {
    const uint32_t lane = lane_id();
    const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / 64);
    for (uint64_t w = (uint64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; w < nwords; w += waves) {
        const uint64_t i = w * 64 + lane;
        int c = 4;
        if (i < T) c = nt4((unsigned char)seq[i]);
        const bool ok = i < T && c < 4;
        const uint64_t lo = __ballot(ok && (c & 1));
        const uint64_t hi = __ballot(ok && (c & 2));
        const uint64_t m = __ballot(!ok);  // past the end reads as N
        if (lane == 0) {
            nmask[w] = m;
            pk[2 * w] = spread32((uint32_t)lo) | (spread32((uint32_t)hi) << 1);
        } else if (lane == 1) {
            pk[2 * w + 1] = spread32((uint32_t)(lo >> 32)) | (spread32((uint32_t)(hi >> 32)) << 1);
        }
    }
}

// 2. canonical s-mer hash per s-mer start (0 where it holds a non-ACGT base:
// such windows are invalid)
__global__ void k_ref_smer(const uint64_t *__restrict__ pk, const uint64_t *__restrict__ nmask, uint64_t ns,
                           SeedParams P, uint64_t *__restrict__ sh) {
    GRID_STRIDE(j, ns) {
        uint64_t h = 0;
        if (!any_n(nmask, j, j + (uint64_t)P.s - 1)) h = xxh64_u64(canonical_code(pk, j, P.s, P.smask));
        sh[j] = h;
    }
}

// 3. window states that need no history
__global__ void k_ref_anchor(const uint64_t *__restrict__ nmask, const uint64_t *__restrict__ sh, uint64_t nk,
                             SeedParams P, uint8_t *__restrict__ st) {
    GRID_STRIDE(p, nk) {
        uint8_t s;
        if (any_n(nmask, p, p + (uint64_t)P.k - 1)) {
            s = ST_INVALID;
        } else {
            const bool seg_first = p == 0 || ((nmask[(p - 1) >> 6] >> ((p - 1) & 63)) & 1);
            s = P.w == 5 ? anchor_state_w<5>(sh, p, seg_first) : anchor_state(sh + p, 0, P.w, seg_first);
        }
        st[p] = s;
    }
}

// 4. each run of tied windows replayed from the known window before it (a
// run never starts a segment: a segment's first window is an anchor). The
// resolved states go to st2, so that st, which decides where runs start,
// never changes under the other lanes.
__global__ void k_ref_resolve(const uint64_t *__restrict__ sh, const uint8_t *__restrict__ st, uint64_t nk,
                              SeedParams P, uint8_t *__restrict__ st2) {
    GRID_STRIDE(p, nk) {
        if (p == 0 || st[p] != ST_UNKNOWN || st[p - 1] == ST_UNKNOWN) continue;
        uint8_t r = st[p - 1];
        for (uint64_t x = p; x < nk && st[x] == ST_UNKNOWN; x++) {
            r = slide(r, sh + (x - 1), 1, P.w);
            st2[x] = r;
        }
    }
}

__device__ __forceinline__ bool is_syncmer(const uint8_t *st, const uint8_t *st2, uint64_t p, int t) {
    uint8_t s = st[p];
    if (s == ST_UNKNOWN) s = st2[p];
    return s == (uint8_t)(t - 1);
}

// 5a. syncmers per tile of TILE windows
__global__ __launch_bounds__(TILE_THREADS) void k_ref_sync_count(const uint8_t *__restrict__ st,
                                                                 const uint8_t *__restrict__ st2, uint64_t nk,
                                                                 int t, uint32_t *__restrict__ cnt) {
    __shared__ uint32_t part[TILE_THREADS / 64];
    const uint64_t base = (uint64_t)blockIdx.x * TILE + (uint64_t)threadIdx.x * TILE_WIN;
    uint32_t c = 0;
    for (int u = 0; u < TILE_WIN; u++) {
        const uint64_t p = base + u;
        if (p < nk && is_syncmer(st, st2, p, t)) c++;
    }
    c = wave_sum(c);
    if (lane_id() == 0) part[threadIdx.x / 64] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int i = 0; i < TILE_THREADS / 64; i++) s += part[i];
        cnt[blockIdx.x] = s;
    }
}

// 5b. syncmer positions in order: tile offset + rank inside the tile
__global__ __launch_bounds__(TILE_THREADS) void k_ref_sync_write(const uint8_t *__restrict__ st,
                                                                 const uint8_t *__restrict__ st2, uint64_t nk,
                                                                 int t, const uint64_t *__restrict__ tile_off,
                                                                 uint64_t *__restrict__ sy_pos) {
    __shared__ uint32_t part[TILE_THREADS / 64];
    const uint64_t base = (uint64_t)blockIdx.x * TILE + (uint64_t)threadIdx.x * TILE_WIN;
    uint32_t flags = 0;
    for (int u = 0; u < TILE_WIN; u++) {
        const uint64_t p = base + u;
        if (p < nk && is_syncmer(st, st2, p, t)) flags |= 1u << u;
    }
    const uint32_t c = (uint32_t)__popc(flags);
    const uint32_t ex = wave_excl_scan(c);
    const uint32_t wtot = wave_sum(c);
    const uint32_t wid = threadIdx.x / 64;
    if (lane_id() == 0) part[wid] = wtot;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t i = 0; i < wid; i++) before += part[i];
    uint64_t at = tile_off[blockIdx.x] + before + ex;
    for (int u = 0; u < TILE_WIN; u++)
        if ((flags >> u) & 1u) sy_pos[at++] = base + u;
}

// 5c. canonical k-mer hash of each syncmer
__global__ void k_ref_sync_hash(const uint64_t *__restrict__ pk, const uint64_t *__restrict__ sy_pos, uint64_t n,
                                SeedParams P, uint64_t *__restrict__ sy_hash) {
    GRID_STRIDE(i, n) sy_hash[i] = xxh64_u64(canonical_code(pk, sy_pos[i], P.k, P.kmask));
}

// 6a. first syncmer of each contig (and n at the end): lower bound of the
// contig's first base among the syncmer positions
__global__ void k_ref_contig_syncmers(const uint64_t *__restrict__ sy_pos, uint64_t n,
                                      const uint64_t *__restrict__ cstart, uint32_t n_refs,
                                      uint64_t *__restrict__ sy_start) {
    GRID_STRIDE(c, (uint64_t)n_refs + 1) {
        const uint64_t key = cstart[c];
        uint64_t lo = 0, cnt = n;
        while (cnt > 0) {
            const uint64_t step = cnt >> 1, it = lo + step;
            if (sy_pos[it] < key) {
                lo = it + 1;
                cnt -= step + 1;
            } else {
                cnt = step;
            }
        }
        sy_start[c] = c == n_refs ? n : lo;
    }
}

// 6b. the randstrobe of each syncmer that starts one
__global__ void k_ref_randstrobes(const uint64_t *__restrict__ sy_pos, const uint64_t *__restrict__ sy_hash,
                                  uint64_t n, const uint64_t *__restrict__ sy_start,
                                  const uint64_t *__restrict__ rs_start, const uint64_t *__restrict__ cstart,
                                  const uint8_t *__restrict__ is_short, uint32_t n_refs, SeedParams P,
                                  uint4 *__restrict__ out) {
    GRID_STRIDE(i, n) {
        // the contig: last c with sy_start[c] <= i
        uint32_t lo = 0, hi = n_refs;  // sy_start[lo] <= i < sy_start[hi]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (sy_start[mid] <= i) lo = mid;
            else hi = mid;
        }
        while (lo + 1 < n_refs && sy_start[lo + 1] <= i) lo++;
        const uint32_t c = lo;
        const uint64_t first = sy_start[c], last = sy_start[c + 1] - 1;  // the contig's syncmers
        if (i + P.w_min > last) continue;  // fewer than w_min syncmers follow
        const uint64_t o = rs_start[c] + (i - first);
        if (is_short[c]) {
            out[o] = make_uint4(0u, 0u, 0u, 0u);
            continue;
        }
        const uint64_t hi_ = sy_hash[i], pi = sy_pos[i];
        const uint64_t wend = std::min<uint64_t>(i + P.w_max, last);
        const uint64_t maxp = pi + P.max_dist_u;
        uint64_t best = i, bh = hi_;
        uint32_t minv = 0xffffffffu;
        for (uint64_t j = i + P.w_min; j <= wend; j++) {
            const uint64_t pj = sy_pos[j];
            if (pj > maxp) break;
            const uint64_t hj = sy_hash[j];
            const uint32_t cnt = (uint32_t)__popcll((hi_ ^ hj) & P.q);
            if (cnt < minv) {
                minv = cnt;
                best = j;
                bh = hj;
            }
        }
        const uint64_t hash = hi_ + bh;
        const uint32_t pos = (uint32_t)(pi - cstart[c]);
        const uint32_t packed = (c << 8) + (uint32_t)(sy_pos[best] - pi);
        out[o] = make_uint4((uint32_t)hash, (uint32_t)(hash >> 32), pos, packed);
    }
}

}  // namespace

void device_seed_reference(const char *const *seqs, const uint64_t *lens, uint32_t n_refs, const gis_params &gp,
                           uint4 **out, uint64_t *n_out, hipStream_t st) {
    *out = nullptr;
    *n_out = 0;
    const SeedParams P = make_seed_params(gp);
    // contig c at cstart[c], one 'N' after each; cstart[n_refs] = T
    std::vector<uint64_t> cstart(n_refs + 1);
    uint64_t T = 0;
    for (uint32_t c = 0; c < n_refs; c++) {
        cstart[c] = T;
        T += lens[c] + 1;
    }
    cstart[n_refs] = T;
    const uint64_t nwords = (T + 63) / 64;
    Dev d_seq, d_pk, d_nm, d_sh, d_st, d_st2, d_cnt, d_toff, d_tmp, d_pos, d_hash, d_cs, d_sys, d_rss, d_short;
    char *seq = dalloc<char>(d_seq, T);
    for (uint32_t c = 0; c < n_refs; c++) {
        if (lens[c]) check(hipMemcpyAsync(seq + cstart[c], seqs[c], lens[c], hipMemcpyHostToDevice, st), "genome upload");
        check(hipMemsetAsync(seq + cstart[c] + lens[c], 'N', 1, st), "contig separator");
    }
    // two pad words: canonical_code reads the word after a window's first
    uint64_t *pk = dalloc<uint64_t>(d_pk, 2 * nwords + 2);
    uint64_t *nm = dalloc<uint64_t>(d_nm, nwords + 1);
    check(hipMemsetAsync(pk + 2 * nwords, 0, 16, st), "pad");
    check(hipMemsetAsync(nm + nwords, 0xff, 8, st), "pad");
    hipLaunchKernelGGL(k_ref_pack, dim3(grid(nwords, 4)), dim3(256), 0, st, seq, T, nwords, pk, nm);
    check(hipGetLastError(), "k_ref_pack");
    check(hipStreamSynchronize(st), "k_ref_pack");
    (void)hipFree(d_seq.p);
    d_seq.p = nullptr;

    const uint64_t ns = T >= (uint64_t)P.s ? T - P.s + 1 : 0;
    const uint64_t nk = T >= (uint64_t)P.k ? T - P.k + 1 : 0;
    if (nk == 0) return;
    uint64_t *sh = dalloc<uint64_t>(d_sh, ns);
    uint8_t *stt = dalloc<uint8_t>(d_st, nk);
    uint8_t *st2 = dalloc<uint8_t>(d_st2, nk);
    hipLaunchKernelGGL(k_ref_smer, dim3(grid(ns, 256)), dim3(256), 0, st, pk, nm, ns, P, sh);
    check(hipGetLastError(), "k_ref_smer");
    hipLaunchKernelGGL(k_ref_anchor, dim3(grid(nk, 256)), dim3(256), 0, st, nm, sh, nk, P, stt);
    check(hipGetLastError(), "k_ref_anchor");
    hipLaunchKernelGGL(k_ref_resolve, dim3(grid(nk, 256)), dim3(256), 0, st, sh, stt, nk, P, st2);
    check(hipGetLastError(), "k_ref_resolve");
    check(hipStreamSynchronize(st), "window states");
    (void)hipFree(d_sh.p);
    d_sh.p = nullptr;

    const uint64_t ntiles = (nk + TILE - 1) / TILE;
    uint32_t *cnt = dalloc<uint32_t>(d_cnt, ntiles);
    uint64_t *toff = dalloc<uint64_t>(d_toff, ntiles + 1);
    uint64_t *tmp = dalloc<uint64_t>(d_tmp, scan_tmp_words(ntiles) + 8);
    hipLaunchKernelGGL(k_ref_sync_count, dim3((unsigned)ntiles), dim3(TILE_THREADS), 0, st, stt, st2, nk, P.t, cnt);
    check(hipGetLastError(), "k_ref_sync_count");
    launch_scan(cnt, toff, ntiles, tmp, st);
    check(hipGetLastError(), "syncmer scan");
    uint64_t nsy = 0;
    check(hipMemcpyAsync(&nsy, toff + ntiles, 8, hipMemcpyDeviceToHost, st), "syncmer count");
    check(hipStreamSynchronize(st), "syncmer count");
    uint64_t *pos = dalloc<uint64_t>(d_pos, nsy);
    hipLaunchKernelGGL(k_ref_sync_write, dim3((unsigned)ntiles), dim3(TILE_THREADS), 0, st, stt, st2, nk, P.t, toff,
                       pos);
    check(hipGetLastError(), "k_ref_sync_write");
    uint64_t *hsh = dalloc<uint64_t>(d_hash, nsy);
    hipLaunchKernelGGL(k_ref_sync_hash, dim3(grid(nsy, 256)), dim3(256), 0, st, pk, pos, nsy, P, hsh);
    check(hipGetLastError(), "k_ref_sync_hash");

    // per contig: syncmer range, randstrobe slots, short contigs
    uint64_t *cs = dalloc<uint64_t>(d_cs, n_refs + 1);
    uint64_t *sys = dalloc<uint64_t>(d_sys, n_refs + 1);
    check(hipMemcpyAsync(cs, cstart.data(), 8ull * (n_refs + 1), hipMemcpyHostToDevice, st), "contig starts");
    hipLaunchKernelGGL(k_ref_contig_syncmers, dim3(grid(n_refs + 1, 256)), dim3(256), 0, st, pos, nsy, cs, n_refs,
                       sys);
    check(hipGetLastError(), "k_ref_contig_syncmers");
    std::vector<uint64_t> sy_start(n_refs + 1), rs_start(n_refs + 1);
    check(hipMemcpyAsync(sy_start.data(), sys, 8ull * (n_refs + 1), hipMemcpyDeviceToHost, st), "syncmer ranges");
    check(hipStreamSynchronize(st), "syncmer ranges");
    std::vector<uint8_t> is_short(std::max<uint32_t>(n_refs, 1), 0);
    uint64_t total = 0;
    for (uint32_t c = 0; c < n_refs; c++) {
        rs_start[c] = total;
        const uint64_t nc = sy_start[c + 1] - sy_start[c];
        total += nc > gp.w_min ? nc - gp.w_min : 0;
        is_short[c] = lens[c] < gp.w_max ? 1 : 0;
    }
    rs_start[n_refs] = total;
    uint64_t *rss = dalloc<uint64_t>(d_rss, n_refs + 1);
    uint8_t *shrt = dalloc<uint8_t>(d_short, n_refs);
    check(hipMemcpyAsync(rss, rs_start.data(), 8ull * (n_refs + 1), hipMemcpyHostToDevice, st), "slot starts");
    check(hipMemcpyAsync(shrt, is_short.data(), n_refs, hipMemcpyHostToDevice, st), "short contigs");
    uint4 *recs = nullptr;
    check(hipMalloc(&recs, std::max<uint64_t>(total, 1) * sizeof(uint4)), "hipMalloc (records)");
    hipLaunchKernelGGL(k_ref_randstrobes, dim3(grid(nsy, 256)), dim3(256), 0, st, pos, hsh, nsy, sys, rss, cs, shrt,
                       n_refs, P, recs);
    const hipError_t e = hipGetLastError();
    const hipError_t e2 = hipStreamSynchronize(st);
    if (e != hipSuccess || e2 != hipSuccess) {
        (void)hipFree(recs);
        throw BuildError{e != hipSuccess ? e : e2, "k_ref_randstrobes"};
    }
    *out = recs;
    *n_out = total;
}

}  // namespace gis
