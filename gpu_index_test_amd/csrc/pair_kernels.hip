// gfx950 kernels of the NAM consumer for paired reads.
//
//  k_pair_order   one wave per NAM list: each NAM's position in the reference
//                 order is the number of NAMs of the list with a smaller key,
//                 counted against the other lanes' keys (lists of up to 64)
//                 or against 64-key tiles in LDS. Keys are unique (each
//                 NAM's (is_rc, ref_id, query_start, ref_start) is its own
//                 first hit), so counting ranks gives the one sorted order the
//                 reference's quicksort reaches. Lists in canonical order are
//                 grouped by (is_rc, ref_id), so a 64-NAM chunk only counts
//                 against the groups it spans.
//  k_pair_type    one lane per pair: type, top_dropoff, MAPQ, per-block type
//                 counts, the bound on its joint entries.
//  k_type_scatter pair ids grouped by type in id order (the host loop of
//                 step2/main.cu:4641-4646) from the scanned block counts.
//  k_joint_wave   one wave per pair with NAMs on both mates:
//                 get_best_scoring_nam_pairs (step2/main.cu:1719-1830),
//                 including its quicksort, whose order of equal entries is
//                 part of the result; k_joint (one lane per pair) takes the
//                 pairs whose entries exceed the wave kernel's LDS.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "kernels.h"
#include "pairs.h"
#include "wave.h"

namespace gis {

// key of a NAM in the reference order: (is_rc, ref_id) as one ordered word,
// the diagonal max(0, ref_start - query_start) for mate 2 (0 for mate 1),
// n_hits descending, query_end, query_start, ref_end, ref_start
struct NamKey {
    uint64_t g;
    int32_t d, nh, qe, qs, re, rs;
};

__device__ __forceinline__ uint64_t group_word(const gis_nam &n) {
    return (uint64_t)n.is_rc << 32 | (uint32_t)(n.ref_id ^ (int32_t)0x80000000);
}

__device__ __forceinline__ NamKey nam_key(const gis_nam &n, bool mate2) {
    NamKey k;
    k.g = group_word(n);
    k.d = mate2 ? max(0, n.ref_start - n.query_start) : 0;
    k.nh = n.n_hits;
    k.qe = n.query_end;
    k.qs = n.query_start;
    k.re = n.ref_end;
    k.rs = n.ref_start;
    return k;
}

__device__ __forceinline__ bool key_less(const NamKey &a, const NamKey &b) {
    if (a.g != b.g) return a.g < b.g;
    if (a.d != b.d) return a.d < b.d;
    if (a.nh != b.nh) return a.nh > b.nh;
    if (a.qe != b.qe) return a.qe < b.qe;
    if (a.qs != b.qs) return a.qs < b.qs;
    if (a.re != b.re) return a.re < b.re;
    return a.rs < b.rs;
}

__device__ __forceinline__ bool key_eq(const NamKey &a, const NamKey &b) {
    return a.g == b.g && a.d == b.d && a.nh == b.nh && a.qe == b.qe && a.qs == b.qs && a.re == b.re &&
           a.rs == b.rs;
}

constexpr int ORDER_WPB = 4;

// gpu_sort_nams, step2/main.cu:3449-3472 (mate 1: step2/main.cu:2275-2286;
// mate 2: Nam::operator<, step2/my_struct.hpp:54-68)
__global__ __launch_bounds__(64 * ORDER_WPB) void k_pair_order(PairDev P) {
    __shared__ NamKey lk[ORDER_WPB][64];
    const uint32_t wid = threadIdx.x >> 6, lane = lane_id();
    const uint32_t n_lists = 2 * P.n_pairs;
    for (uint32_t r = blockIdx.x * ORDER_WPB + wid; r < n_lists; r += gridDim.x * ORDER_WPB) {
        const uint64_t o = P.off[r];
        const uint32_t cnt = (uint32_t)(P.off[r + 1] - o);
        const bool mate2 = r >= P.n_pairs;
        const gis_nam *in = P.in + o;
        gis_nam *out = P.ord + o;
        if (cnt <= 64) {
            // the whole list in the lanes: each lane ranks its NAM against
            // every other one read straight from the lanes
            const bool valid = lane < cnt;
            gis_nam mine{};
            NamKey ki{};
            if (valid) {
                mine = in[lane];
                ki = nam_key(mine, mate2);
            }
            const uint32_t glo = (uint32_t)ki.g, ghi = (uint32_t)(ki.g >> 32);
            uint32_t rank = 0;
            bool dup = false;
            for (uint32_t q = 0; q < cnt; q++) {
                NamKey kj;
                kj.g = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)ghi, (int)q) << 32 |
                       (uint32_t)__builtin_amdgcn_readlane((int)glo, (int)q);
                kj.d = __builtin_amdgcn_readlane(ki.d, (int)q);
                kj.nh = __builtin_amdgcn_readlane(ki.nh, (int)q);
                kj.qe = __builtin_amdgcn_readlane(ki.qe, (int)q);
                kj.qs = __builtin_amdgcn_readlane(ki.qs, (int)q);
                kj.re = __builtin_amdgcn_readlane(ki.re, (int)q);
                kj.rs = __builtin_amdgcn_readlane(ki.rs, (int)q);
                rank += key_less(kj, ki) ? 1u : 0u;
                dup |= q != lane && key_eq(kj, ki);
            }
            if (valid) out[rank] = mine;
            if (__ballot(valid && dup) != 0 && lane == 0) atomicOr(P.err, PAIR_ERR_DUPLICATE);
            continue;
        }
        for (uint32_t i0 = 0; i0 < cnt; i0 += 64) {
            const uint32_t i = i0 + lane;
            const bool valid = i < cnt;
            gis_nam mine{};
            NamKey ki{};
            if (valid) {
                mine = in[i];
                ki = nam_key(mine, mate2);
            }
            uint32_t u0 = 0, u1 = cnt;
            if (P.grouped) {
                // the groups this chunk spans: [first index of the first NAM's
                // group, one past the last NAM's group)
                const uint32_t last = min(i0 + 63, cnt - 1);
                const uint64_t gf = group_word(in[i0]), gl = group_word(in[last]);
                uint32_t lo = 0, hi = i0;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (group_word(in[mid]) < gf) lo = mid + 1;
                    else hi = mid;
                }
                u0 = lo;
                lo = last + 1;
                hi = cnt;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (group_word(in[mid]) <= gl) lo = mid + 1;
                    else hi = mid;
                }
                u1 = lo;
            }
            uint32_t rank = 0;
            bool dup = false;
            for (uint32_t t = u0; t < u1; t += 64) {
                const uint32_t j = t + lane;
                if (j < u1) lk[wid][lane] = nam_key(in[j], mate2);
                wave_sync();
                const uint32_t m = min(64u, u1 - t);
                if (valid) {
                    for (uint32_t q = 0; q < m; q++) {
                        const NamKey kj = lk[wid][q];
                        rank += key_less(kj, ki) ? 1u : 0u;
                        dup |= (t + q != i) && key_eq(kj, ki);
                    }
                }
                wave_sync();
            }
            if (valid) out[u0 + rank] = mine;
            if (dup) atomicOr(P.err, PAIR_ERR_DUPLICATE);
        }
    }
}

// top_dropoff, step2/main.cu:889-898 (0 for an empty list, which the
// reference never asks for)
__device__ __forceinline__ float top_dropoff(const gis_nam *L, uint32_t c) {
    if (!c) return 0.0f;
    const int nh0 = L[0].n_hits;
    if (nh0 <= 2) return 1.0f;
    if (c > 1) return (float)L[1].n_hits / (float)nh0;
    return 0.0f;
}

// float to int, truncating and saturating (NaN -> 0)
__device__ __forceinline__ int f2i_sat(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return (int)0x80000000;
    return (int)x;
}

// get_mapq(nams, nams[0]), step2/main.cu:900-910; the natural log is the
// correctly rounded float logarithm (log in double rounded to float), and
// the uint8 result wraps a negative value as the reference's return type does
__device__ __forceinline__ uint8_t mapq_of(const gis_nam *L, uint32_t c) {
    if (c <= 1) return 60;
    const float s1 = L[0].score, s2 = L[1].score;
    const double mm = L[0].n_hits / 10.0;
    const float min_matches = (float)(mm < 1.0 ? mm : 1.0);
    const float lg = (float)log((double)s1);
    const float a = 40.0f * (1.0f - __fdiv_rn(s2, s1));
    const float v = __fmul_rn(__fmul_rn(a, min_matches), lg);
    const int u = f2i_sat(v);
    return (uint8_t)(u < 60 ? u : 60);
}

struct PairNam {
    int32_t ref_id, rs, qs, nh;
    uint8_t is_rc;
};

__device__ __forceinline__ PairNam pair_nam(const gis_nam &n) {
    return PairNam{n.ref_id, n.ref_start, n.query_start, n.n_hits, n.is_rc};
}

// is_proper_nam_pair, step2/main.cu:870-887; mu + 10 sigma is one fused
// multiply-add, as the reference's device code computes it
__device__ __forceinline__ bool proper_pair(const PairNam &a, const PairNam &b, float lim) {
    if (a.ref_id != b.ref_id || a.is_rc == b.is_rc) return false;
    const int x = max(0, a.rs - a.qs);
    const int y = max(0, b.rs - b.qs);
    const bool r1_r2 = b.is_rc && x <= y && (float)(y - x) < lim;
    const bool r2_r1 = a.is_rc && y <= x && (float)(x - y) < lim;
    return r1_r2 || r2_r1;
}

__device__ __forceinline__ uint64_t joint_bound(uint32_t c1, uint32_t c2, int mt) {
    if (!c1 || !c2) return 0;
    const uint64_t prod = (uint64_t)c1 * c2;
    const uint64_t first = prod < (uint64_t)(2 * mt + 1) ? prod : (uint64_t)(2 * mt + 1);
    return first + min((uint32_t)mt, c1) + min((uint32_t)mt, c2);
}

// gpu_pre_cal_type, step2/main.cu:3552-3581, with the fresh insert-size
// estimate's mu and sigma
__global__ __launch_bounds__(PAIR_BLOCK) void k_pair_type(PairDev P) {
    __shared__ uint32_t cnt[5];
    if (threadIdx.x < 5) cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t p = blockIdx.x * PAIR_BLOCK + threadIdx.x;
    const uint32_t n = P.n_pairs;
    if (p < n) {
        const uint64_t o1 = P.off[p], o2 = P.off[n + p];
        const uint32_t c1 = (uint32_t)(P.off[p + 1] - o1), c2 = (uint32_t)(P.off[n + p + 1] - o2);
        const gis_nam *A = P.ord + o1, *B = P.ord + o2;
        const float d1 = top_dropoff(A, c1), d2 = top_dropoff(B, c2);
        uint8_t t;
        if (!c1 && !c2) t = 0;
        else if (c1 && !c2) t = 1;
        else if (!c1 && c2) t = 2;
        else if (d1 < P.dropoff && d2 < P.dropoff &&
                 proper_pair(pair_nam(A[0]), pair_nam(B[0]), __builtin_fmaf(10.0f, P.sigma, P.mu)))
            t = 3;
        else t = 4;
        P.types[p] = t;
        P.drop[p] = d1;
        P.drop[n + p] = d2;
        P.mapq[p] = mapq_of(A, c1);
        P.mapq[n + p] = mapq_of(B, c2);
        if (P.joint) P.jbound[p] = (uint32_t)joint_bound(c1, c2, P.max_tries);
        atomicAdd(&cnt[t], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 5) P.tcnt[threadIdx.x * P.nblk + blockIdx.x] = cnt[threadIdx.x];
}

__global__ __launch_bounds__(PAIR_BLOCK) void k_type_scatter(PairDev P) {
    constexpr int W = PAIR_BLOCK / 64;
    __shared__ uint32_t wc[W][5];
    const uint32_t p = blockIdx.x * PAIR_BLOCK + threadIdx.x;
    const uint32_t wid = threadIdx.x >> 6;
    const int t = p < P.n_pairs ? (int)P.types[p] : -1;
    uint32_t wrank = 0;
    for (int tt = 0; tt < 5; tt++) {
        const uint64_t m = __ballot(t == tt);
        if (t == tt) wrank = mbcnt(m);
        if (lane_id() == 0) wc[wid][tt] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (t >= 0) {
        uint32_t before = 0;
        for (uint32_t w = 0; w < wid; w++) before += wc[w][t];
        const uint64_t pos = P.tpos[(uint64_t)t * P.nblk + blockIdx.x] + before + wrank;
        P.tids[pos] = p;
    }
}

__global__ void k_type_totals(PairDev P) {
    const uint32_t t = threadIdx.x;
    if (t <= 5) P.ttot[t] = P.tpos[(uint64_t)t * P.nblk];
}

// an entry's NAM fields that the NamPair comparator reads; an absent NAM is
// Nam{} with ref_start = -1
struct CmpNam {
    float score;
    int32_t qe, qs, re, rs;
    uint8_t is_rc;
};

__device__ __forceinline__ CmpNam cmp_nam(const gis_nam *L, int32_t i) {
    if (i < 0) return CmpNam{0.0f, 0, 0, 0, -1, 0};
    const gis_nam &n = L[i];
    return CmpNam{n.score, n.query_end, n.query_start, n.ref_end, n.ref_start, n.is_rc};
}

// the comparator of step2/main.cu:1790-1828 (strict)
__device__ __forceinline__ bool entry_less(const gis_nam_pair &x, const gis_nam_pair &y, const gis_nam *A,
                                           const gis_nam *B) {
    if (x.score != y.score) return x.score > y.score;
    {
        const CmpNam a = cmp_nam(A, x.i1), b = cmp_nam(A, y.i1);
        if (a.score != b.score) return a.score > b.score;
        if (a.is_rc != b.is_rc) return !a.is_rc;
        if (a.qe != b.qe) return a.qe < b.qe;
        if (a.qs != b.qs) return a.qs < b.qs;
        if (a.re != b.re) return a.re < b.re;
        if (a.rs != b.rs) return a.rs < b.rs;
    }
    const CmpNam a = cmp_nam(B, x.i2), b = cmp_nam(B, y.i2);
    if (a.score != b.score) return a.score > b.score;
    if (a.is_rc != b.is_rc) return !a.is_rc;
    if (a.qe != b.qe) return a.qe < b.qe;
    if (a.qs != b.qs) return a.qs < b.qs;
    if (a.re != b.re) return a.re < b.re;
    return a.rs < b.rs;
}

// quick_sort_iterative, step2/main.cu:317-359: Lomuto partitions around the
// last element. Partitions of disjoint ranges do not interact, so taking the
// smaller part first (the larger one waits on a stack of at most log2 n
// ranges) ends in the same array as the reference's order of ranges.
__device__ void lomuto_sort(gis_nam_pair *E, int n, const gis_nam *A, const gis_nam *B) {
    int st_lo[24], st_hi[24];
    int sp = 0;
    int lo = 0, hi = n - 1;
    for (;;) {
        while (lo < hi) {
            const gis_nam_pair pivot = E[hi];
            int i = lo - 1;
            for (int j = lo; j < hi; j++) {
                const gis_nam_pair ej = E[j];
                if (entry_less(ej, pivot, A, B)) {
                    ++i;
                    E[j] = E[i];
                    E[i] = ej;
                }
            }
            const int p = i + 1;
            E[hi] = E[p];
            E[p] = pivot;
            // ranges of one element are left as they are, like the reference
            const int lsz = p - lo, rsz = hi - p;
            if (lsz <= rsz) {
                if (rsz >= 2) {
                    st_lo[sp] = p + 1;
                    st_hi[sp] = hi;
                    sp++;
                }
                hi = p - 1;
            } else {
                if (lsz >= 2) {
                    st_lo[sp] = lo;
                    st_hi[sp] = p - 1;
                    sp++;
                }
                lo = p + 1;
            }
        }
        if (sp == 0) break;
        sp--;
        lo = st_lo[sp];
        hi = st_hi[sp];
    }
}

constexpr int FLAG_WORDS = PAIR_MAX_TRIES / 32;
constexpr uint32_t JOINT_LDS_CAP = 128;  // entries a pair may have in k_joint_wave
constexpr int FLAG_STRIDE = 2 * FLAG_WORDS + 1;  // odd stride: lanes fall on different banks

// get_best_scoring_nam_pairs, step2/main.cu:1719-1830
__global__ __launch_bounds__(PAIR_BLOCK) void k_joint(PairDev P) {
    __shared__ uint32_t flags[PAIR_BLOCK * FLAG_STRIDE];
    const uint32_t p = blockIdx.x * PAIR_BLOCK + threadIdx.x;
    const uint32_t n = P.n_pairs;
    if (p >= n) return;
    const uint32_t cap = P.jbound[p];
    // bounds within JOINT_LDS_CAP (and pairs without entries) are k_joint_wave's
    if (cap <= JOINT_LDS_CAP) return;
    uint32_t *f1 = flags + threadIdx.x * FLAG_STRIDE, *f2 = f1 + FLAG_WORDS;
    for (int w = 0; w < FLAG_WORDS; w++) f1[w] = f2[w] = 0;
    const uint64_t o1 = P.off[p], o2 = P.off[n + p];
    const uint32_t c1 = (uint32_t)(P.off[p + 1] - o1), c2 = (uint32_t)(P.off[n + p + 1] - o2);
    const gis_nam *A = P.ord + o1, *B = P.ord + o2;
    gis_nam_pair *E = P.jtmp + P.jboff[p];
    const int mt = P.max_tries;
    const float lim = __builtin_fmaf(10.0f, P.sigma, P.mu);
    uint32_t size = 0;
    bool over = false;
    int best = 0;
    auto push = [&](int sc, int a, int b) {
        if (size < cap) E[size] = gis_nam_pair{sc, a, b};
        else over = true;
        size++;
    };
    for (uint32_t i = 0; i < c1; i++) {
        const PairNam a = pair_nam(A[i]);
        for (uint32_t j = 0; j < c2; j++) {
            const PairNam b = pair_nam(B[j]);
            const int joint = a.nh + b.nh;
            // joint < 0.5 * best in double is 2 joint < best in integers
            if (2 * joint < best || size > (uint32_t)(2 * mt)) break;
            if (proper_pair(a, b, lim)) {
                push(joint, (int)i, (int)j);
                if (i < (uint32_t)mt) f1[i >> 5] |= 1u << (i & 31);
                if (j < (uint32_t)mt) f2[j >> 5] |= 1u << (j & 31);
                best = max(best, joint);
            }
        }
        if (size > (uint32_t)(2 * mt)) break;
    }
    const int best1 = best > 0 ? best : A[0].n_hits;
    const uint32_t lim1 = min(c1, (uint32_t)mt);
    for (uint32_t i = 0; i < lim1; i++) {
        const int nh = A[i].n_hits;
        if (nh < best1 / 2) break;
        if (f1[i >> 5] >> (i & 31) & 1u) continue;
        push(nh, (int)i, -1);
    }
    const int best2 = best > 0 ? best : B[0].n_hits;
    const uint32_t lim2 = min(c2, (uint32_t)mt);
    for (uint32_t i = 0; i < lim2; i++) {
        const int nh = B[i].n_hits;
        if (nh < best2 / 2) break;
        if (f2[i >> 5] >> (i & 31) & 1u) continue;
        push(nh, -1, (int)i);
    }
    if (over) {
        atomicOr(P.err, PAIR_ERR_JOINT);
        P.jcnt[p] = 0;
        return;
    }
    lomuto_sort(E, (int)size, A, B);
    P.jcnt[p] = size;
}

// ---------------------------------------------------------------------------
// get_best_scoring_nam_pairs with one wave per pair (pairs whose entry bound
// fits JOINT_LDS_CAP; max_tries up to 31 always does)
//
// The reference's scan over (i, j) is sequential: at each j it stops when
// the joint hits fall below half the best pushed so far or more than
// 2 max_tries entries are pushed, then pushes (i, j) if the NAMs form a
// proper pair. Over 64 lanes j at once, the running best and count only
// change at proper lanes, so the row splits into segments ending at each
// proper lane: one ballot tests a segment's lanes against the running
// values, and if none stops, the proper lane ending it pushes.
// The entries are then sorted by ranking keys in LDS. Without two equal keys
// the sorted order is unique and equals the reference quicksort's; with
// equal keys one lane replays the quicksort on the entry indices.

constexpr int JOINT_WPB = 4;

// the NamPair comparator's fields of one entry
struct EntryKey {
    int32_t sc;
    float s1;
    int32_t qe1, qs1, re1, rs1;
    uint32_t rc1;
    float s2;
    int32_t qe2, qs2, re2, rs2;
    uint32_t rc2;
};

__device__ __forceinline__ EntryKey entry_key(const gis_nam_pair &e, const gis_nam *A, const gis_nam *B) {
    const CmpNam a = cmp_nam(A, e.i1), b = cmp_nam(B, e.i2);
    return EntryKey{e.score, a.score, a.qe, a.qs, a.re, a.rs, a.is_rc, b.score, b.qe, b.qs, b.re, b.rs, b.is_rc};
}

// entry_less on materialised keys
__device__ __forceinline__ bool ekey_less(const EntryKey &x, const EntryKey &y) {
    if (x.sc != y.sc) return x.sc > y.sc;
    if (x.s1 != y.s1) return x.s1 > y.s1;
    if (x.rc1 != y.rc1) return !x.rc1;
    if (x.qe1 != y.qe1) return x.qe1 < y.qe1;
    if (x.qs1 != y.qs1) return x.qs1 < y.qs1;
    if (x.re1 != y.re1) return x.re1 < y.re1;
    if (x.rs1 != y.rs1) return x.rs1 < y.rs1;
    if (x.s2 != y.s2) return x.s2 > y.s2;
    if (x.rc2 != y.rc2) return !x.rc2;
    if (x.qe2 != y.qe2) return x.qe2 < y.qe2;
    if (x.qs2 != y.qs2) return x.qs2 < y.qs2;
    if (x.re2 != y.re2) return x.re2 < y.re2;
    return x.rs2 < y.rs2;
}

__device__ __forceinline__ bool ekey_eq(const EntryKey &x, const EntryKey &y) {
    return x.sc == y.sc && x.s1 == y.s1 && x.rc1 == y.rc1 && x.qe1 == y.qe1 && x.qs1 == y.qs1 &&
           x.re1 == y.re1 && x.rs1 == y.rs1 && x.s2 == y.s2 && x.rc2 == y.rc2 && x.qe2 == y.qe2 &&
           x.qs2 == y.qs2 && x.re2 == y.re2 && x.rs2 == y.rs2;
}

// quick_sort_iterative (step2/main.cu:317-359) on entry indices, smaller
// part first as in lomuto_sort
__device__ void lomuto_idx(uint8_t *I, int n, const EntryKey *K) {
    int st_lo[16], st_hi[16];
    int sp = 0;
    int lo = 0, hi = n - 1;
    for (;;) {
        while (lo < hi) {
            const uint8_t pv = I[hi];
            int i = lo - 1;
            for (int j = lo; j < hi; j++) {
                const uint8_t ej = I[j];
                if (ekey_less(K[ej], K[pv])) {
                    ++i;
                    I[j] = I[i];
                    I[i] = ej;
                }
            }
            const int p = i + 1;
            I[hi] = I[p];
            I[p] = pv;
            const int lsz = p - lo, rsz = hi - p;
            if (lsz <= rsz) {
                if (rsz >= 2) {
                    st_lo[sp] = p + 1;
                    st_hi[sp] = hi;
                    sp++;
                }
                hi = p - 1;
            } else {
                if (lsz >= 2) {
                    st_lo[sp] = lo;
                    st_hi[sp] = p - 1;
                    sp++;
                }
                lo = p + 1;
            }
        }
        if (sp == 0) break;
        sp--;
        lo = st_lo[sp];
        hi = st_hi[sp];
    }
}

__global__ __launch_bounds__(64 * JOINT_WPB) void k_joint_wave(PairDev P) {
    __shared__ gis_nam_pair lE[JOINT_WPB][JOINT_LDS_CAP];
    __shared__ EntryKey lK[JOINT_WPB][JOINT_LDS_CAP];
    __shared__ uint32_t lF[JOINT_WPB][2 * FLAG_WORDS];
    __shared__ uint8_t lI[JOINT_WPB][JOINT_LDS_CAP];
    const uint32_t wid = threadIdx.x >> 6, lane = lane_id();
    gis_nam_pair *E = lE[wid];
    EntryKey *K = lK[wid];
    uint32_t *f1 = lF[wid], *f2 = f1 + FLAG_WORDS;
    uint8_t *I = lI[wid];
    const uint32_t n = P.n_pairs;
    const int mt = P.max_tries;
    const float lim = __builtin_fmaf(10.0f, P.sigma, P.mu);
    for (uint32_t p = blockIdx.x * JOINT_WPB + wid; p < n; p += gridDim.x * JOINT_WPB) {
        const uint32_t cap = P.jbound[p];
        if (cap == 0 || cap > JOINT_LDS_CAP) {
            // no entries, or a bound beyond LDS: k_joint takes the pair
            if (cap == 0 && lane == 0) P.jcnt[p] = 0;
            continue;
        }
        if (lane < 2 * FLAG_WORDS) f1[lane] = 0;
        wave_sync();
        const uint64_t o1 = P.off[p], o2 = P.off[n + p];
        const uint32_t c1 = (uint32_t)(P.off[p + 1] - o1), c2 = (uint32_t)(P.off[n + p + 1] - o2);
        const gis_nam *A = P.ord + o1, *B = P.ord + o2;
        // the first 64 NAMs of each list stay in registers (lane i holds
        // NAM i); rows and columns beyond them are loaded as they come
        PairNam ra{}, rb{};
        if (lane < c1) ra = pair_nam(A[lane]);
        if (lane < c2) rb = pair_nam(B[lane]);
        uint32_t size = 0;
        int best = 0;
        for (uint32_t i = 0; i < c1; i++) {
            PairNam a;
            if (i < 64) {
                a.ref_id = __builtin_amdgcn_readlane(ra.ref_id, (int)i);
                a.rs = __builtin_amdgcn_readlane(ra.rs, (int)i);
                a.qs = __builtin_amdgcn_readlane(ra.qs, (int)i);
                a.nh = __builtin_amdgcn_readlane(ra.nh, (int)i);
                a.is_rc = (uint8_t)__builtin_amdgcn_readlane((int)ra.is_rc, (int)i);
            } else {
                a = pair_nam(A[i]);
            }
            bool stopped = false;
            for (uint32_t jc = 0; jc < c2 && !stopped; jc += 64) {
                const uint32_t j = jc + lane;
                const bool valid = j < c2;
                PairNam b = rb;
                if (jc && valid) b = pair_nam(B[j]);
                const int joint = a.nh + b.nh;
                const bool prop = valid && proper_pair(a, b, lim);
                // the lanes in order, one segment per proper lane: the stop
                // test (joint < 0.5 * best in double is 2 joint < best in
                // integers) runs on the segment's lanes with the running
                // values, then the proper lane that ends it pushes
                uint64_t rem = __ballot(prop);
                uint32_t from = 0;
                for (;;) {
                    const uint32_t pl = rem ? (uint32_t)__builtin_ctzll(rem) : 64u;
                    const uint32_t to = pl < 64 ? pl : 63u;
                    const uint64_t seg = (to == 63 ? ~0ull : ((1ull << (to + 1)) - 1)) & (~0ull << from);
                    const bool stop = valid && (2 * joint < best || size > (uint32_t)(2 * mt));
                    const uint64_t sm = __ballot(stop) & seg;
                    if (sm) {
                        stopped = true;
                        break;
                    }
                    if (pl == 64) break;
                    const int jp = __builtin_amdgcn_readlane(joint, (int)pl);
                    if (lane == pl) {
                        E[size] = gis_nam_pair{jp, (int)i, (int)j};
                        if (i < (uint32_t)mt) f1[i >> 5] |= 1u << (i & 31);
                        if (j < (uint32_t)mt) f2[j >> 5] |= 1u << (j & 31);
                    }
                    size++;
                    best = max(best, jp);
                    rem &= rem - 1;
                    from = pl + 1;
                    if (from >= 64) break;
                }
            }
            if (size > (uint32_t)(2 * mt)) break;
        }
        wave_sync();
        // NAMs of each mate in no pushed pair, while their hits reach half
        // the best joint hits (the list's first NAM's hits without one)
        for (int m = 0; m < 2; m++) {
            const gis_nam *L = m ? B : A;
            const uint32_t c = m ? c2 : c1;
            const uint32_t *fl = m ? f2 : f1;
            const int bestm = best > 0 ? best : __builtin_amdgcn_readfirstlane(m ? rb.nh : ra.nh);
            const uint32_t lm = min(c, (uint32_t)mt);
            for (uint32_t ic = 0; ic < lm; ic += 64) {
                const uint32_t i = ic + lane;
                const bool valid = i < lm;
                const int nh = !valid ? 0 : (ic == 0 ? (m ? rb.nh : ra.nh) : L[i].n_hits);
                const bool stop = valid && nh < bestm / 2;
                const uint64_t sm = __ballot(stop);
                const uint32_t first = sm ? (uint32_t)__builtin_ctzll(sm) : 64u;
                const bool push = valid && lane < first && !((fl[i >> 5] >> (i & 31)) & 1u);
                const uint64_t qm = __ballot(push);
                if (push) E[size + mbcnt(qm)] = m ? gis_nam_pair{nh, -1, (int)i} : gis_nam_pair{nh, (int)i, -1};
                size += (uint32_t)__popcll(qm);
                if (first < 64) break;
            }
        }
        wave_sync();
        const uint32_t ne = size;
        for (uint32_t e = lane; e < ne; e += 64) K[e] = entry_key(E[e], A, B);
        wave_sync();
        uint32_t rank[2] = {0, 0};
        bool tie = false;
        for (int k = 0; k < 2; k++) {
            const uint32_t e = lane + 64 * k;
            if (e < ne) {
                const EntryKey me = K[e];
                uint32_t r = 0;
                for (uint32_t f = 0; f < ne; f++) {
                    const EntryKey o = K[f];
                    if (ekey_less(o, me)) r++;
                    else if (f != e && ekey_eq(o, me)) tie = true;
                }
                rank[k] = r;
            }
        }
        gis_nam_pair *out = P.jtmp + P.jboff[p];
        if (!__any(tie)) {
            for (int k = 0; k < 2; k++) {
                const uint32_t e = lane + 64 * k;
                if (e < ne) out[rank[k]] = E[e];
            }
        } else {
            if (lane == 0) {
                for (uint32_t e = 0; e < ne; e++) I[e] = (uint8_t)e;
                lomuto_idx(I, (int)ne, K);
            }
            wave_sync();
            for (uint32_t e = lane; e < ne; e += 64) out[e] = E[I[e]];
        }
        if (lane == 0) P.jcnt[p] = ne;
        wave_sync();
    }
}

__global__ __launch_bounds__(PAIR_BLOCK) void k_joint_compact(PairDev P) {
    const uint32_t p = blockIdx.x * PAIR_BLOCK + threadIdx.x;
    if (p >= P.n_pairs) return;
    const uint32_t c = P.jcnt[p];
    const gis_nam_pair *src = P.jtmp + P.jboff[p];
    gis_nam_pair *dst = P.jout + P.joff[p];
    for (uint32_t k = 0; k < c; k++) dst[k] = src[k];
}

size_t pair_blocks(uint32_t n_pairs) { return ((size_t)n_pairs + PAIR_BLOCK - 1) / PAIR_BLOCK; }

void launch_pairs_classify(const PairDev &P, uint64_t *scan_tmp, hipStream_t st) {
    const uint32_t n_lists = 2 * P.n_pairs;
    if (n_lists) {
        const unsigned g = (unsigned)std::min<uint64_t>(((uint64_t)n_lists + ORDER_WPB - 1) / ORDER_WPB, 65536);
        hipLaunchKernelGGL(k_pair_order, dim3(g), dim3(64 * ORDER_WPB), 0, st, P);
    }
    const unsigned nb = (unsigned)pair_blocks(P.n_pairs);
    if (nb) hipLaunchKernelGGL(k_pair_type, dim3(nb), dim3(PAIR_BLOCK), 0, st, P);
    launch_scan(P.tcnt, const_cast<uint64_t *>(P.tpos), 5ull * nb, scan_tmp, st);
    if (nb) hipLaunchKernelGGL(k_type_scatter, dim3(nb), dim3(PAIR_BLOCK), 0, st, P);
    hipLaunchKernelGGL(k_type_totals, dim3(1), dim3(64), 0, st, P);
    if (P.joint) launch_scan(P.jbound, const_cast<uint64_t *>(P.jboff), P.n_pairs, scan_tmp, st);
}

void launch_pairs_joint(const PairDev &P, uint64_t *scan_tmp, hipStream_t st) {
    const unsigned nb = (unsigned)pair_blocks(P.n_pairs);
    if (P.n_pairs) {
        const unsigned g =
            (unsigned)std::min<uint64_t>(((uint64_t)P.n_pairs + JOINT_WPB - 1) / JOINT_WPB, 65536);
        hipLaunchKernelGGL(k_joint_wave, dim3(g), dim3(64 * JOINT_WPB), 0, st, P);
    }
    // bounds beyond LDS need max_tries above 31
    if (nb && 4u * (uint32_t)P.max_tries + 1 > JOINT_LDS_CAP)
        hipLaunchKernelGGL(k_joint, dim3(nb), dim3(PAIR_BLOCK), 0, st, P);
    launch_scan(P.jcnt, const_cast<uint64_t *>(P.joff), P.n_pairs, scan_tmp, st);
    if (nb) hipLaunchKernelGGL(k_joint_compact, dim3(nb), dim3(PAIR_BLOCK), 0, st, P);
}

}  // namespace gis
