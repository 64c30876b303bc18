/*
 * CPU oracle for strobealign's seeding path -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain C restatement of the reference CPU algorithm; it is the checker the
 * HIP product is compared against and the timed CPU baseline. The product
 * library never links or calls this file. See oracle.h for the API.
 *
 * Reference citations are file:line relative to the reference repository.
 *
 * Deliberate, documented differences from the reference text:
 *  - hits_per_ref is a robin_hood map whose iteration order is unspecified
 *    (step2/nam.cu:368); this oracle iterates refs in ascending ref_id. That
 *    only changes the append order of NAMs, which orc_canonicalize removes.
 *  - Index entries with equal (hash, position) are ordered by `packed` (the
 *    reference's pdqsort leaves them unspecified, step2/index.cu:170).
 *  - The equal-hash run walk stops at the end of the table; the reference
 *    would read past it if the last hash were 0xFFFF...FF
 *    (step2/nam.cu:53, step2/index.hpp:83-89).
 *  - k = 32 uses an all-ones k-mer mask; the reference shifts 1ULL by 64
 *    there (step2/randstrobes.hpp:332), which is undefined behaviour.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* xxh64 reduced to one u64 input, step2/hash.hpp:106-119. Constants are the
 * published XXH64 primes (step2/hash.hpp:97-101). */
#define P1 0x9E3779B185EBCA87ULL
#define P2 0xC2B2AE3D27D4EB4FULL
#define P3 0x165667B19E3779F9ULL
#define P4 0x85EBCA77C2B2AE63ULL
#define P5 0x27D4EB2F165667C5ULL

static inline uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

uint64_t orc_xxh64(uint64_t input) {
    uint64_t result = P5 + 8;
    input *= P2;
    input = rotl64(input, 31);
    result ^= input * P1;
    result = rotl64(result, 27);
    result = result * P1 + P4;
    result ^= result >> 33;
    result *= P2;
    result ^= result >> 29;
    result *= P3;
    result ^= result >> 32;
    return result;
}

/* seq_nt4_table, step2/randstrobes.hpp:22-43: A/a->0 C/c->1 G/g->2 T/t/U/u->3 */
int orc_nt4(unsigned char c) {
    switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': case 'U': case 'u': return 3;
    default: return 4;
    }
}

/* ------------------------------------------------------------------ */
/* IndexParameters constructor + verify, step2/indexparameters.hpp:16-34,
 * 47-69, 80-85. */
int orc_params_make(int canonical_read_length, int k, int s, int l, int u, int q,
                    int max_dist, orc_params *out) {
    if (k <= 7 || k > 32) return ORC_BAD_PARAMETER;
    if (s > k) return ORC_BAD_PARAMETER;
    if ((k - s) % 2 != 0) return ORC_BAD_PARAMETER;
    int w_min_i = k / (k - s + 1) + l;
    unsigned w_min = (unsigned)(w_min_i > 0 ? w_min_i : 0);
    unsigned w_max = (unsigned)(k / (k - s + 1) + u);
    if (max_dist > 255) return ORC_BAD_PARAMETER;
    if (w_min > w_max) return ORC_BAD_PARAMETER;
    out->canonical_read_length = canonical_read_length;
    out->k = k;
    out->s = s;
    out->t_syncmer = (k - s) / 2 + 1;
    out->l = l;
    out->u = u;
    out->q = (uint64_t)(int64_t)q; /* int q stored in a uint64_t member */
    out->max_dist = max_dist;
    out->w_min = w_min;
    out->w_max = w_max;
    return ORC_OK;
}

/* Profile table and from_read_length, step2/indexparameters.cpp:25-83 */
typedef struct { int canonical_read_length, r_threshold, k, s_offset, l, u; } orc_profile;
static const orc_profile PROFILES[] = {
    {50, 90, 20, -4, -3, 2},   {100, 110, 20, -4, -2, 2}, {125, 135, 20, -4, -1, 4},
    {150, 175, 20, -4, 1, 7},  {250, 275, 20, -4, 4, 13}, {300, 375, 22, -4, 2, 12},
    {400, 0x7fffffff, 23, -6, 2, 12},
};

int orc_params_from_read_length(int read_length, int k, int s, int l, int u, int c,
                                int max_seed_len, orc_params *out) {
    const int default_c = 8;
    int canonical_read_length = 50;
    for (size_t i = 0; i < sizeof(PROFILES) / sizeof(PROFILES[0]); i++) {
        const orc_profile *p = &PROFILES[i];
        if (read_length <= p->r_threshold) {
            if (k == ORC_DEFAULT) k = p->k;
            if (s == ORC_DEFAULT) s = k + p->s_offset;
            if (l == ORC_DEFAULT) l = p->l;
            if (u == ORC_DEFAULT) u = p->u;
            canonical_read_length = p->canonical_read_length;
            break;
        }
    }
    int max_dist;
    if (max_seed_len == ORC_DEFAULT) {
        max_dist = canonical_read_length - 70 > k ? canonical_read_length - 70 : k;
        if (max_dist > 255) max_dist = 255;
    } else {
        max_dist = max_seed_len - k;
    }
    /* int q = std::pow(2, c) - 1 (double pow, converted to int) */
    int q = (int)(pow(2.0, (double)(c == ORC_DEFAULT ? default_c : c)) - 1.0);
    return orc_params_make(canonical_read_length, k, s, l, u, q, max_dist, out);
}

/* ------------------------------------------------------------------ */
/* SyncmerIterator::next, step2/randstrobes.cu:18-79, run to completion
 * (canonical_syncmers, step2/randstrobes.cu:83-94). The std::deque of s-mer
 * hashes is a ring buffer holding at most k-s+2 entries. */
size_t orc_syncmers(const char *seq, size_t len, const orc_params *p,
                    orc_syncmer *out, size_t cap) {
    const size_t k = (size_t)p->k, s = (size_t)p->s, t = (size_t)p->t_syncmer;
    const uint64_t kmask = k >= 32 ? ~0ULL : (1ULL << (2 * k)) - 1;
    const uint64_t smask = s >= 32 ? ~0ULL : (1ULL << (2 * s)) - 1;
    const uint64_t kshift = (k - 1) * 2, sshift = (s - 1) * 2;
    uint64_t ring[40];
    size_t qhead = 0, qsize = 0; /* deque: element j is ring[(qhead+j)%40] */
    uint64_t qs_min_val = UINT64_MAX;
    size_t qs_min_pos = (size_t)-1;
    size_t l = 0;
    uint64_t xk0 = 0, xk1 = 0, xs0 = 0, xs1 = 0;
    size_t n = 0;
#define QAT(j) ring[(qhead + (j)) % 40]
    for (size_t i = 0; i < len; ++i) {
        int c = orc_nt4((unsigned char)seq[i]);
        if (c < 4) {
            xk0 = (xk0 << 2 | (uint64_t)c) & kmask;
            xk1 = xk1 >> 2 | (uint64_t)(3 - c) << kshift;
            xs0 = (xs0 << 2 | (uint64_t)c) & smask;
            xs1 = xs1 >> 2 | (uint64_t)(3 - c) << sshift;
            if (++l < s) continue;
            uint64_t ys = xs0 < xs1 ? xs0 : xs1;
            uint64_t hash_s = orc_xxh64(ys);
            QAT(qsize) = hash_s;
            qsize++;
            if (qsize < k - s + 1) continue;
            if (qsize == k - s + 1) {
                for (size_t j = 0; j < qsize; j++) {
                    if (QAT(j) < qs_min_val) {
                        qs_min_val = QAT(j);
                        qs_min_pos = i - k + j + 1;
                    }
                }
            } else {
                qhead = (qhead + 1) % 40; /* pop_front */
                qsize--;
                if (qs_min_pos == i - k) {
                    qs_min_val = UINT64_MAX;
                    qs_min_pos = i - s + 1;
                    for (long j = (long)qsize - 1; j >= 0; j--) {
                        if (QAT(j) < qs_min_val) {
                            qs_min_val = QAT(j);
                            qs_min_pos = i - k + (size_t)j + 1;
                        }
                    }
                } else if (hash_s < qs_min_val) {
                    qs_min_val = hash_s;
                    qs_min_pos = i - s + 1;
                }
            }
            if (qs_min_pos == i - k + t) {
                uint64_t yk = xk0 < xk1 ? xk0 : xk1;
                if (n < cap) {
                    out[n].hash = orc_xxh64(yk);
                    out[n].position = i - k + 1;
                }
                n++;
            }
        } else {
            qs_min_val = UINT64_MAX;
            qs_min_pos = (size_t)-1;
            l = 0;
            xs0 = xs1 = xk0 = xk1 = 0;
            qhead = 0;
            qsize = 0;
        }
    }
#undef QAT
    return n;
}

/* RandstrobeIterator::get, step2/randstrobes.cu:136-159 */
static void randstrobe_get(const orc_syncmer *sy, size_t n, unsigned idx, const orc_params *p,
                           uint64_t *hash, uint32_t *pos1, uint32_t *pos2) {
    size_t we = (size_t)idx + p->w_max;
    unsigned w_end = (unsigned)(we < n - 1 ? we : n - 1);
    orc_syncmer strobe1 = sy[idx];
    size_t max_position = strobe1.position + (unsigned)p->max_dist;
    unsigned w_start = idx + p->w_min;
    uint64_t min_val = UINT64_MAX;
    orc_syncmer strobe2 = strobe1;
    for (unsigned i = w_start; i <= w_end && sy[i].position <= max_position; i++) {
        uint64_t res = (uint64_t)__builtin_popcountll((strobe1.hash ^ sy[i].hash) & p->q);
        if (res < min_val) {
            min_val = res;
            strobe2 = sy[i];
        }
    }
    *hash = strobe1.hash + strobe2.hash; /* randstrobe_hash, step2/randstrobes.hpp:61-63 */
    *pos1 = (uint32_t)strobe1.position;
    *pos2 = (uint32_t)strobe2.position;
}

/* RandstrobeIterator loop over a syncmer list, emitting QueryRandstrobes
 * (step2/randstrobes.cu:208-216 and 231-239) */
static size_t emit_query_randstrobes(const orc_syncmer *sy, size_t n, const orc_params *p,
                                     uint8_t is_reverse, orc_qrs *out, size_t m, size_t cap) {
    for (unsigned idx = 0; idx + p->w_min < n; idx++) {
        uint64_t h;
        uint32_t p1, p2;
        randstrobe_get(sy, n, idx, p, &h, &p1, &p2);
        if (m < cap) {
            memset(&out[m], 0, sizeof(orc_qrs));
            out[m].hash = h;
            out[m].start = p1;
            out[m].end = p2 + (uint32_t)p->k;
            out[m].is_reverse = is_reverse;
        }
        m++;
    }
    return m;
}

/* randstrobes_query, step2/randstrobes.cu:195-241 */
size_t orc_randstrobes_query(const char *seq, size_t len, const orc_params *p,
                             orc_qrs *out, size_t cap) {
    if (len < p->w_max) return 0;
    size_t scap = len + 1;
    orc_syncmer *sy = (orc_syncmer *)malloc(scap * sizeof(orc_syncmer));
    if (!sy) return 0;
    size_t n = orc_syncmers(seq, len, p, sy, scap);
    if (n == 0) { free(sy); return 0; }
    size_t m = emit_query_randstrobes(sy, n, p, 0, out, 0, cap);
    /* reverse the syncmers and mirror positions, step2/randstrobes.cu:221-224 */
    for (size_t i = 0; i < n / 2; i++) {
        orc_syncmer tmp = sy[i];
        sy[i] = sy[n - 1 - i];
        sy[n - 1 - i] = tmp;
    }
    for (size_t i = 0; i < n; i++) sy[i].position = len - sy[i].position - (size_t)p->k;
    m = emit_query_randstrobes(sy, n, p, 1, out, m, cap);
    free(sy);
    return m;
}

/* ------------------------------------------------------------------ */
/* StrobemerIndex::pick_bits, step2/index.cu:137-141 */
int orc_pick_bits(uint64_t total_length, const orc_params *p) {
    uint64_t est = total_length / (uint64_t)(p->k - p->s + 1);
    if (est == 0) return 8;
    int b = (int)log2((double)est) - 1;
    if (b < 8) b = 8;
    if (b > 31) b = 31;
    return b;
}

/* LSD radix sort of 16-byte records on the key (hash, position, packed);
 * passes whose digit is constant across all records are skipped. */
static int radix_sort_ref(orc_ref_randstrobe *a, uint64_t n) {
    if (n < 2) return ORC_OK;
    orc_ref_randstrobe *tmp = (orc_ref_randstrobe *)malloc(n * sizeof(*a));
    if (!tmp) return ORC_NOMEM;
    orc_ref_randstrobe *src = a, *dst = tmp;
    /* key bytes from least significant: packed(4), position(4), hash(8) */
    for (int pass = 0; pass < 16; pass++) {
        uint64_t cnt[256];
        memset(cnt, 0, sizeof(cnt));
        for (uint64_t i = 0; i < n; i++) {
            unsigned d;
            if (pass < 4) d = (src[i].packed >> (8 * pass)) & 0xff;
            else if (pass < 8) d = (src[i].position >> (8 * (pass - 4))) & 0xff;
            else d = (unsigned)((src[i].hash >> (8 * (pass - 8))) & 0xff);
            cnt[d]++;
        }
        int trivial = 0;
        for (int d = 0; d < 256; d++) if (cnt[d] == n) trivial = 1;
        if (trivial) continue;
        uint64_t sum = 0;
        for (int d = 0; d < 256; d++) { uint64_t c = cnt[d]; cnt[d] = sum; sum += c; }
        for (uint64_t i = 0; i < n; i++) {
            unsigned d;
            if (pass < 4) d = (src[i].packed >> (8 * pass)) & 0xff;
            else if (pass < 8) d = (src[i].position >> (8 * (pass - 4))) & 0xff;
            else d = (unsigned)((src[i].hash >> (8 * (pass - 8))) & 0xff);
            dst[cnt[d]++] = src[i];
        }
        orc_ref_randstrobe *t = src; src = dst; dst = t;
    }
    if (src != a) memcpy(a, src, n * sizeof(*a));
    free(tmp);
    return ORC_OK;
}

typedef struct {
    const char *const *seqs;
    const uint64_t *lens;
    uint32_t n_refs;
    const orc_params *p;
    uint64_t *counts;        /* per ref */
    uint64_t *offsets;       /* per ref */
    orc_ref_randstrobe *out;
    atomic_uint next;
    int phase;
} build_job;

/* count_randstrobes / assign_randstrobes, step2/index.cu:28-40,280-311:
 * randstrobes of one contig are the forward randstrobes of its syncmers,
 * packed = (ref_index << 8) + (strobe2_pos - strobe1_pos). */
static void *build_worker(void *arg) {
    build_job *J = (build_job *)arg;
    for (;;) {
        unsigned j = atomic_fetch_add(&J->next, 1);
        if (j >= J->n_refs) break;
        const char *seq = J->seqs[j];
        size_t len = (size_t)J->lens[j];
        size_t scap = len + 1;
        orc_syncmer *sy = (orc_syncmer *)malloc(scap * sizeof(orc_syncmer));
        size_t n = sy ? orc_syncmers(seq, len, J->p, sy, scap) : 0;
        if (J->phase == 0) {
            J->counts[j] = n < J->p->w_min ? 0 : n - J->p->w_min;
        } else if (len >= J->p->w_max) {
            uint64_t off = J->offsets[j];
            for (unsigned idx = 0; idx + J->p->w_min < n; idx++) {
                uint64_t h;
                uint32_t p1, p2;
                randstrobe_get(sy, n, idx, J->p, &h, &p1, &p2);
                uint32_t packed = (uint32_t)j << 8;
                packed = packed + (p2 - p1);
                J->out[off].hash = h;
                J->out[off].position = p1;
                J->out[off].packed = packed;
                off++;
            }
        }
        free(sy);
    }
    return NULL;
}

static int cmp_u64_desc(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x > y ? -1 : (x < y ? 1 : 0);
}

/* StrobemerIndex::populate, step2/index.cu:143-244 */
int orc_index_build(const char *const *seqs, const uint64_t *lens, uint32_t n_refs,
                    const orc_params *p, int bits, float f, int n_threads,
                    orc_index **out) {
    uint64_t total_length = 0;
    for (uint32_t j = 0; j < n_refs; j++) total_length += lens[j];
    if (bits == -1) bits = orc_pick_bits(total_length, p);
    if (bits < 8 || bits > 31) return ORC_BAD_PARAMETER;
    if (n_threads < 1) n_threads = 1;

    build_job J;
    J.seqs = seqs; J.lens = lens; J.n_refs = n_refs; J.p = p;
    J.counts = (uint64_t *)calloc(n_refs + 1, sizeof(uint64_t));
    J.offsets = (uint64_t *)calloc(n_refs + 1, sizeof(uint64_t));
    J.out = NULL;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)n_threads);
    for (int phase = 0; phase < 2; phase++) {
        J.phase = phase;
        atomic_store(&J.next, 0);
        if (phase == 1) {
            uint64_t tot = 0;
            for (uint32_t j = 0; j < n_refs; j++) { J.offsets[j] = tot; tot += J.counts[j]; }
            J.out = (orc_ref_randstrobe *)malloc((tot ? tot : 1) * sizeof(orc_ref_randstrobe));
            if (!J.out) return ORC_NOMEM;
        }
        for (int t = 0; t < n_threads; t++) pthread_create(&th[t], NULL, build_worker, &J);
        for (int t = 0; t < n_threads; t++) pthread_join(th[t], NULL);
    }
    free(th);
    uint64_t n = 0;
    for (uint32_t j = 0; j < n_refs; j++) n += J.counts[j];
    free(J.counts);
    free(J.offsets);
    orc_ref_randstrobe *rs = J.out;

    int rc = radix_sort_ref(rs, n);
    if (rc) return rc;

    orc_index *idx = (orc_index *)calloc(1, sizeof(orc_index));
    idx->randstrobes = rs;
    idx->n = n;
    idx->bits = bits;
    idx->params = *p;
    idx->owns = 1;
    uint64_t nb = (1ULL << bits) + 1;
    idx->starts = (uint64_t *)malloc(nb * sizeof(uint64_t));
    uint64_t ns = 0;

    /* bucket starts + repetition statistics, step2/index.cu:176-240 */
    uint64_t *counts = (uint64_t *)malloc((n + 2) * sizeof(uint64_t));
    uint64_t n_counts = 0;
    uint64_t unique_mers = n == 0 ? 0 : 1;
    uint64_t prev_hash = n == 0 ? 0 : rs[0].hash;
    unsigned count = 0;
    for (uint64_t position = 0; position < n; ++position) {
        uint64_t cur = rs[position].hash;
        if (cur == prev_hash) { ++count; continue; }
        ++unique_mers;
        if (count == 1) {
            ++idx->tot_occur_once;
        } else {
            if (count > 100) ++idx->tot_high_ab; else ++idx->tot_mid_ab;
            counts[n_counts++] = count;
        }
        count = 1;
        unsigned cur_n = (unsigned)(cur >> (64 - bits));
        while (ns <= cur_n) idx->starts[ns++] = position;
        prev_hash = cur;
    }
    if (count == 1) {
        ++idx->tot_occur_once;
    } else {
        if (count > 100) ++idx->tot_high_ab; else ++idx->tot_mid_ab;
        counts[n_counts++] = count;
    }
    while (ns < nb) idx->starts[ns++] = n;
    qsort(counts, n_counts, sizeof(uint64_t), cmp_u64_desc);
    uint64_t index_cutoff = (uint64_t)((float)unique_mers * f);
    unsigned filter_cutoff;
    if (n_counts > 0) {
        filter_cutoff = (unsigned)(index_cutoff < n_counts ? counts[index_cutoff] : counts[n_counts - 1]);
        if (filter_cutoff < 30U) filter_cutoff = 30U;
        if (filter_cutoff > 100U) filter_cutoff = 100U;
    } else {
        filter_cutoff = 30;
    }
    free(counts);
    idx->filter_cutoff = filter_cutoff;
    idx->distinct_strobemers = unique_mers;
    idx->index_cutoff = index_cutoff;
    *out = idx;
    return ORC_OK;
}

int orc_index_wrap(orc_ref_randstrobe *rs, uint64_t n, uint64_t *starts, int bits,
                   uint32_t filter_cutoff, const orc_params *p, orc_index **out) {
    if (bits < 8 || bits > 31) return ORC_BAD_PARAMETER;
    orc_index *idx = (orc_index *)calloc(1, sizeof(orc_index));
    idx->randstrobes = rs;
    idx->n = n;
    idx->starts = starts;
    idx->bits = bits;
    idx->filter_cutoff = filter_cutoff;
    idx->params = *p;
    idx->owns = 0;
    *out = idx;
    return ORC_OK;
}

void orc_index_free(orc_index *idx) {
    if (!idx) return;
    if (idx->owns) {
        free(idx->randstrobes);
        free(idx->starts);
    }
    free(idx);
}

/* StrobemerIndex::find, step2/index.hpp:57-81 (the unrolled copy at
 * step2/nam.cu:571-669 computes the same value). */
uint64_t orc_find(const orc_index *idx, uint64_t key) {
    const unsigned top_n = (unsigned)(key >> (64 - idx->bits));
    uint64_t ps = idx->starts[top_n];
    uint64_t pe = idx->starts[top_n + 1];
    if (ps == pe) return UINT64_MAX;
    if (pe - ps < 4) {
        for (; ps < pe; ++ps) {
            if (idx->randstrobes[ps].hash == key) return ps;
            if (idx->randstrobes[ps].hash > key) return UINT64_MAX;
        }
        return UINT64_MAX;
    }
    /* std::lower_bound on hash */
    uint64_t lo = ps, cnt = pe - ps;
    while (cnt > 0) {
        uint64_t step = cnt / 2, it = lo + step;
        if (idx->randstrobes[it].hash < key) { lo = it + 1; cnt -= step + 1; }
        else cnt = step;
    }
    if (lo < idx->n && idx->randstrobes[lo].hash == key) return lo;
    return UINT64_MAX;
}

/* get_hash / is_filtered, step2/index.hpp:83-93 */
static inline uint64_t get_hash(const orc_index *idx, uint64_t position) {
    return position < idx->n ? idx->randstrobes[position].hash : UINT64_MAX;
}

int orc_is_filtered(const orc_index *idx, uint64_t position) {
    return get_hash(idx, position) == get_hash(idx, position + idx->filter_cutoff);
}

/* get_count, step2/index.hpp:115-147 */
uint32_t orc_get_count(const orc_index *idx, uint64_t position) {
    const uint64_t key = idx->randstrobes[position].hash;
    const unsigned top_n = (unsigned)(key >> (64 - idx->bits));
    uint64_t pe = idx->starts[top_n + 1];
    uint64_t count = 1;
    if (pe - position < 8) {
        for (uint64_t q = position + 1; q < pe; ++q) {
            if (idx->randstrobes[q].hash == key) count += 1;
            else break;
        }
        return (uint32_t)count;
    }
    /* std::upper_bound on hash over [position, pe) */
    uint64_t lo = position, cnt = pe - position;
    while (cnt > 0) {
        uint64_t step = cnt / 2, it = lo + step;
        if (!(key < idx->randstrobes[it].hash)) { lo = it + 1; cnt -= step + 1; }
        else cnt = step;
    }
    return (uint32_t)((lo - 1) - position + 1);
}

/* ------------------------------------------------------------------ */
/* growable (ref_id, hit) list: the per-strand hits_per_ref map in insertion
 * order */
typedef struct {
    int32_t *ref;
    orc_hit *hit;
    size_t n, cap;
} hitlist;

static int hl_push(hitlist *L, int32_t ref, orc_hit h) {
    if (L->n == L->cap) {
        size_t nc = L->cap ? 2 * L->cap : 256;
        int32_t *r = (int32_t *)realloc(L->ref, nc * sizeof(int32_t));
        orc_hit *x = (orc_hit *)realloc(L->hit, nc * sizeof(orc_hit));
        if (!r || !x) return ORC_NOMEM;
        L->ref = r; L->hit = x; L->cap = nc;
    }
    L->ref[L->n] = ref;
    L->hit[L->n] = h;
    L->n++;
    return ORC_OK;
}

static void hl_free(hitlist *L) { free(L->ref); free(L->hit); memset(L, 0, sizeof(*L)); }

/* add_to_hits_per_ref, step2/nam.cu:45-62 */
static void add_to_hits_per_ref(hitlist *L, int query_start, int query_end,
                                const orc_index *idx, uint64_t position) {
    int min_diff = 0x7fffffff;
    const int k = idx->params.k;
    const uint64_t hash = get_hash(idx, position);
    for (; position < idx->n && get_hash(idx, position) == hash; ++position) {
        int ref_start = (int)idx->randstrobes[position].position;
        int ref_end = ref_start + (int)(idx->randstrobes[position].packed & 0xff) + k;
        int d = (query_end - query_start) - (ref_end - ref_start);
        int diff = d < 0 ? -d : d;
        if (diff <= min_diff) {
            orc_hit h = {query_start, query_end, ref_start, ref_end};
            hl_push(L, (int32_t)(idx->randstrobes[position].packed >> 8), h);
            min_diff = diff;
        }
    }
}

/* score of a finished NAM, step2/nam.cu:459-464 */
static void finish_nam(orc_nam *n) {
    int qspan = n->query_end - n->query_start, rspan = n->ref_end - n->ref_start;
    int mx = qspan > rspan ? qspan : rspan, mn = qspan < rspan ? qspan : rspan;
    n->score = (2 * mn - mx) > 0 ? (float)(n->n_hits * (2 * mn - mx)) : 1.0f;
}

typedef struct { orc_nam *v; size_t n, cap; } namvec;

static int nv_push(namvec *V, const orc_nam *x) {
    if (V->n == V->cap) {
        size_t nc = V->cap ? 2 * V->cap : 16;
        orc_nam *p = (orc_nam *)realloc(V->v, nc * sizeof(orc_nam));
        if (!p) return ORC_NOMEM;
        V->v = p; V->cap = nc;
    }
    V->v[V->n++] = *x;
    return ORC_OK;
}

/* Hit::operator<, step2/my_struct.hpp:76-79 */
static int cmp_hit(const void *a, const void *b) {
    const orc_hit *x = (const orc_hit *)a, *y = (const orc_hit *)b;
    if (x->query_start != y->query_start) return x->query_start < y->query_start ? -1 : 1;
    if (x->ref_start != y->ref_start) return x->ref_start < y->ref_start ? -1 : 1;
    return 0;
}

/* merge_hits_into_nams: hit-outer loop for one (strand, ref) group,
 * step2/nam.cu:372-498 */
static void merge_group_hit_outer(orc_hit *hits, size_t nh, int32_t ref_id, int k,
                                  int is_revcomp, int sort, namvec *nams, namvec *open) {
    if (sort) qsort(hits, nh, sizeof(orc_hit), cmp_hit);
    open->n = 0;
    unsigned prev_q_start = 0;
    for (size_t hi = 0; hi < nh; hi++) {
        const orc_hit *h = &hits[hi];
        int is_added = 0;
        for (size_t oi = 0; oi < open->n; oi++) {
            orc_nam *o = &open->v[oi];
            if ((o->query_prev_hit_startpos < h->query_start) && (h->query_start <= o->query_end) &&
                (o->ref_prev_hit_startpos < h->ref_start) && (h->ref_start <= o->ref_end)) {
                if ((h->query_end > o->query_end) && (h->ref_end > o->ref_end)) {
                    o->query_end = h->query_end;
                    o->ref_end = h->ref_end;
                    o->query_prev_hit_startpos = h->query_start;
                    o->ref_prev_hit_startpos = h->ref_start;
                    o->n_hits++;
                    is_added = 1;
                    break;
                } else if ((h->query_end <= o->query_end) && (h->ref_end <= o->ref_end)) {
                    o->query_prev_hit_startpos = h->query_start;
                    o->ref_prev_hit_startpos = h->ref_start;
                    o->n_hits++;
                    is_added = 1;
                    break;
                }
            }
        }
        if (!is_added) {
            orc_nam n;
            memset(&n, 0, sizeof(n));
            n.query_start = h->query_start;
            n.query_end = h->query_end;
            n.ref_start = h->ref_start;
            n.ref_end = h->ref_end;
            n.ref_id = ref_id;
            n.query_prev_hit_startpos = h->query_start;
            n.ref_prev_hit_startpos = h->ref_start;
            n.n_hits = 1;
            n.is_rc = (uint8_t)is_revcomp;
            nv_push(open, &n);
        }
        if ((unsigned)h->query_start > prev_q_start + (unsigned)k) {
            size_t w = 0;
            for (size_t oi = 0; oi < open->n; oi++) {
                orc_nam *n = &open->v[oi];
                if (n->query_end < h->query_start) {
                    finish_nam(n);
                    n->nam_id = (int32_t)nams->n;
                    nv_push(nams, n);
                } else {
                    open->v[w++] = *n; /* std::remove_if keeps order */
                }
            }
            open->n = w;
            prev_q_start = (unsigned)h->query_start;
        }
    }
    for (size_t oi = 0; oi < open->n; oi++) {
        orc_nam *n = &open->v[oi];
        finish_nam(n);
        n->nam_id = (int32_t)nams->n;
        nv_push(nams, n);
    }
}

/* merge_hits_into_nams_fast: NAM-outer loop per query_start group for one
 * (strand, ref) group, step2/nam.cu:115-329 */
static void merge_group_nam_outer(orc_hit *hits, size_t nh, int32_t ref_id, int k,
                                  int is_revcomp, int sort, namvec *nams, namvec *open,
                                  unsigned char *is_added) {
    if (sort) qsort(hits, nh, sizeof(orc_hit), cmp_hit);
    open->n = 0;
    unsigned prev_q_start = 0;
    for (size_t i = 0; i < nh;) {
        size_t i_start = i, i_end = i + 1;
        while (i_end < nh && hits[i_end].query_start == hits[i].query_start) i_end++;
        i = i_end;
        size_t i_size = i_end - i_start;
        memset(is_added, 0, i_size);
        int query_start = hits[i_start].query_start;
        size_t cnt_done = 0;
        qsort(hits + i_start, i_size, sizeof(orc_hit), cmp_hit);
        for (size_t oi = 0; oi < open->n; oi++) {
            orc_nam *o = &open->v[oi];
            /* lower_bound(ref_start < rprev+1) and lower_bound(ref_start < ref_end+1) */
            size_t lower = i_start, upper;
            while (lower < i_end && hits[lower].ref_start < o->ref_prev_hit_startpos + 1) lower++;
            upper = lower;
            while (upper < i_end && hits[upper].ref_start < o->ref_end + 1) upper++;
            for (size_t j = lower; j < upper; j++) {
                if (is_added[j - i_start]) continue;
                if (query_start <= o->query_end) {
                    orc_hit *h = &hits[j];
                    if (o->ref_prev_hit_startpos < h->ref_start && h->ref_start <= o->ref_end) {
                        if ((h->query_end > o->query_end) && (h->ref_end > o->ref_end)) {
                            o->query_end = h->query_end;
                            o->ref_end = h->ref_end;
                            o->query_prev_hit_startpos = h->query_start;
                            o->ref_prev_hit_startpos = h->ref_start;
                            o->n_hits++;
                            is_added[j - i_start] = 1;
                            cnt_done++;
                            break;
                        } else if ((h->query_end <= o->query_end) && (h->ref_end <= o->ref_end)) {
                            o->query_prev_hit_startpos = h->query_start;
                            o->ref_prev_hit_startpos = h->ref_start;
                            o->n_hits++;
                            is_added[j - i_start] = 1;
                            cnt_done++;
                            break;
                        }
                    }
                }
            }
            if (cnt_done == i_size) break;
        }
        for (size_t j = 0; j < i_size; j++) {
            if (!is_added[j]) {
                orc_nam n;
                memset(&n, 0, sizeof(n));
                n.query_start = hits[i_start + j].query_start;
                n.query_end = hits[i_start + j].query_end;
                n.ref_start = hits[i_start + j].ref_start;
                n.ref_end = hits[i_start + j].ref_end;
                n.ref_id = ref_id;
                n.query_prev_hit_startpos = hits[i_start + j].query_start;
                n.ref_prev_hit_startpos = hits[i_start + j].ref_start;
                n.n_hits = 1;
                n.is_rc = (uint8_t)is_revcomp;
                nv_push(open, &n);
            }
        }
        if ((unsigned)query_start > prev_q_start + (unsigned)k) {
            size_t w = 0;
            for (size_t oi = 0; oi < open->n; oi++) {
                orc_nam *n = &open->v[oi];
                if (n->query_end < query_start) {
                    finish_nam(n);
                    n->nam_id = (int32_t)nams->n;
                    nv_push(nams, n);
                } else {
                    open->v[w++] = *n;
                }
            }
            open->n = w;
            prev_q_start = (unsigned)query_start;
        }
    }
    for (size_t oi = 0; oi < open->n; oi++) {
        orc_nam *n = &open->v[oi];
        finish_nam(n);
        n->nam_id = (int32_t)nams->n;
        nv_push(nams, n);
    }
}

typedef struct { int32_t ref; uint32_t ord; } refkey;

static int cmp_refkey(const void *a, const void *b) {
    const refkey *x = (const refkey *)a, *y = (const refkey *)b;
    if (x->ref != y->ref) return x->ref < y->ref ? -1 : 1;
    return x->ord < y->ord ? -1 : (x->ord > y->ord);
}

/* One strand of merge_hits_into_nams_forward_and_reverse{,_fast}
 * (step2/nam.cu:515-539): group the strand's (ref, hit) list by ref keeping
 * insertion order inside a ref, then merge each group. */
static int merge_strand(const hitlist *L, int k, int is_revcomp, int nam_outer, int sort,
                        namvec *nams) {
    size_t nh = L->n;
    if (nh == 0) return ORC_OK;
    refkey *keys = (refkey *)malloc(nh * sizeof(refkey));
    orc_hit *grp = (orc_hit *)malloc(nh * sizeof(orc_hit));
    unsigned char *added = (unsigned char *)malloc(nh);
    namvec open = {0};
    if (!keys || !grp || !added) return ORC_NOMEM;
    for (size_t i = 0; i < nh; i++) { keys[i].ref = L->ref[i]; keys[i].ord = (uint32_t)i; }
    qsort(keys, nh, sizeof(refkey), cmp_refkey);
    for (size_t i = 0; i < nh;) {
        size_t j = i;
        while (j < nh && keys[j].ref == keys[i].ref) { grp[j - i] = L->hit[keys[j].ord]; j++; }
        if (nam_outer)
            merge_group_nam_outer(grp, j - i, keys[i].ref, k, is_revcomp, sort, nams, &open, added);
        else
            merge_group_hit_outer(grp, j - i, keys[i].ref, k, is_revcomp, sort, nams, &open);
        i = j;
    }
    free(open.v);
    free(keys);
    free(grp);
    free(added);
    return ORC_OK;
}

static long emit(const namvec *V, orc_nam *out, size_t cap) {
    if (V->n > cap) return -1;
    if (V->n) memcpy(out, V->v, V->n * sizeof(orc_nam));
    return (long)V->n;
}

long orc_merge_hits(const int32_t *ref_ids, const orc_hit *hits, size_t nh, int k,
                    int is_revcomp, int nam_outer, orc_nam *out, size_t cap) {
    hitlist L = {0};
    for (size_t i = 0; i < nh; i++) hl_push(&L, ref_ids[i], hits[i]);
    namvec V = {0};
    merge_strand(&L, k, is_revcomp, nam_outer, 1, &V);
    long r = emit(&V, out, cap);
    hl_free(&L);
    free(V.v);
    return r;
}

/* find_nams, step2/nam.cu:553-747 */
long orc_find_nams(const orc_index *idx, const orc_qrs *q, size_t nq,
                   orc_nam *out, size_t cap, float *nonrep_fraction, uint64_t *n_hits) {
    hitlist hp[2] = {{0}, {0}};
    int nr_good_hits = 0, total_hits = 0;
    for (size_t i = 0; i < nq; i++) {
        uint64_t position = orc_find(idx, q[i].hash);
        if (position != UINT64_MAX) {
            total_hits++;
            if (orc_is_filtered(idx, position)) continue;
            nr_good_hits++;
            add_to_hits_per_ref(&hp[q[i].is_reverse ? 1 : 0], (int)q[i].start, (int)q[i].end, idx, position);
        }
    }
    float frac = total_hits > 0 ? ((float)nr_good_hits) / ((float)total_hits) : 1.0f;
    if (nonrep_fraction) *nonrep_fraction = frac;
    if (n_hits) *n_hits = hp[0].n + hp[1].n;
    namvec V = {0};
    merge_strand(&hp[0], idx->params.k, 0, 0, 1, &V);
    merge_strand(&hp[1], idx->params.k, 1, 0, 1, &V);
    long r = emit(&V, out, cap);
    hl_free(&hp[0]);
    hl_free(&hp[1]);
    free(V.v);
    return r;
}

/* RescueHit, step2/my_struct.hpp:88-98 */
typedef struct { uint64_t position; uint32_t count, query_start, query_end; } rhit;

/* cmp1: (count, query_start, query_end), step2/nam.cu:915-918 */
static int cmp_rh1(const void *a, const void *b) {
    const rhit *x = (const rhit *)a, *y = (const rhit *)b;
    if (x->count != y->count) return x->count < y->count ? -1 : 1;
    if (x->query_start != y->query_start) return x->query_start < y->query_start ? -1 : 1;
    if (x->query_end != y->query_end) return x->query_end < y->query_end ? -1 : 1;
    return 0;
}

/* cmp2: query_start, step2/nam.cu:920-924 */
static int cmp_rh2(const void *a, const void *b) {
    const rhit *x = (const rhit *)a, *y = (const rhit *)b;
    return x->query_start < y->query_start ? -1 : (x->query_start > y->query_start);
}

/* find_nams_rescue (pre_sort variant), step2/nam.cu:927-984. hard_cutoff is
 * the literal 1000 at step2/nam.cu:964. */
long orc_find_nams_rescue(const orc_index *idx, const orc_qrs *q, size_t nq,
                          unsigned rescue_cutoff, unsigned hard_cutoff,
                          orc_nam *out, size_t cap) {
    rhit *h[2];
    size_t nh[2] = {0, 0};
    h[0] = (rhit *)malloc((nq + 1) * sizeof(rhit));
    h[1] = (rhit *)malloc((nq + 1) * sizeof(rhit));
    for (size_t i = 0; i < nq; i++) {
        uint64_t position = orc_find(idx, q[i].hash);
        if (position != UINT64_MAX) {
            rhit r = {position, orc_get_count(idx, position), q[i].start, q[i].end};
            int s = q[i].is_reverse ? 1 : 0;
            h[s][nh[s]++] = r;
        }
    }
    hitlist hp[2] = {{0}, {0}};
    for (int s = 0; s < 2; s++) {
        qsort(h[s], nh[s], sizeof(rhit), cmp_rh1);
        size_t cnt = 0;
        for (size_t i = 0; i < nh[s]; i++) {
            if ((h[s][i].count > rescue_cutoff && cnt >= 5) || h[s][i].count > hard_cutoff) break;
            cnt++;
        }
        qsort(h[s], cnt, sizeof(rhit), cmp_rh2);
        for (size_t i = 0; i < cnt; i++)
            add_to_hits_per_ref(&hp[s], (int)h[s][i].query_start, (int)h[s][i].query_end, idx,
                                h[s][i].position);
    }
    namvec V = {0};
    merge_strand(&hp[0], idx->params.k, 0, 1, 0, &V);
    merge_strand(&hp[1], idx->params.k, 1, 1, 0, &V);
    long r = emit(&V, out, cap);
    hl_free(&hp[0]);
    hl_free(&hp[1]);
    free(V.v);
    free(h[0]);
    free(h[1]);
    return r;
}

/* canonical NAM order: (is_rc, ref_id, query_start, ref_start); nam_id = rank */
static int cmp_nam_canon(const void *a, const void *b) {
    const orc_nam *x = (const orc_nam *)a, *y = (const orc_nam *)b;
    if (x->is_rc != y->is_rc) return x->is_rc < y->is_rc ? -1 : 1;
    if (x->ref_id != y->ref_id) return x->ref_id < y->ref_id ? -1 : 1;
    if (x->query_start != y->query_start) return x->query_start < y->query_start ? -1 : 1;
    if (x->ref_start != y->ref_start) return x->ref_start < y->ref_start ? -1 : 1;
    return 0;
}

void orc_canonicalize(orc_nam *nams, size_t n) {
    qsort(nams, n, sizeof(orc_nam), cmp_nam_canon);
    for (size_t i = 0; i < n; i++) nams[i].nam_id = (int32_t)i;
}

/* One read through the whole path: randstrobes_query -> find_nams -> the
 * caller's rescue decision (step2/main.cu:3209-3216: nonrepetitive fraction
 * below 0.7, compared in double, or no hits) -> find_nams_rescue. */
long orc_map_read(const orc_index *idx, const char *seq, size_t len,
                  unsigned rescue_cutoff, unsigned hard_cutoff,
                  orc_nam *out, size_t cap, float *frac, uint8_t *flags) {
    size_t qcap = 2 * (len + 1);
    orc_qrs *q = (orc_qrs *)malloc(qcap * sizeof(orc_qrs));
    size_t nq = orc_randstrobes_query(seq, len, &idx->params, q, qcap);
    float fr = 1.0f;
    uint64_t nh = 0;
    long n = orc_find_nams(idx, q, nq, out, cap, &fr, &nh);
    uint8_t fl = 0;
    if ((double)fr < 0.7 || nh == 0) {
        n = orc_find_nams_rescue(idx, q, nq, rescue_cutoff, hard_cutoff, out, cap);
        fl |= ORC_FLAG_RESCUED;
    }
    free(q);
    if (n > 0) orc_canonicalize(out, (size_t)n);
    if (frac) *frac = fr;
    if (flags) *flags = fl;
    return n;
}

typedef struct {
    const orc_index *idx;
    const char *seqs;
    const uint64_t *offsets;
    const uint32_t *lens;
    size_t n_reads;
    unsigned rescue_cutoff, hard_cutoff;
    const uint64_t *nam_offsets;
    orc_nam *nams;
    uint32_t *nam_counts;
    float *fracs;
    uint8_t *flags;
    atomic_size_t next;
    atomic_int err;
} batch_job;

static void *batch_worker(void *arg) {
    batch_job *B = (batch_job *)arg;
    size_t cap = 4096;
    orc_nam *buf = (orc_nam *)malloc(cap * sizeof(orc_nam));
    const size_t chunk = 10000; /* CommandLineOptions::chunk_size, step2/cmdline.hpp:10 */
    for (;;) {
        size_t lo = atomic_fetch_add(&B->next, chunk);
        if (lo >= B->n_reads) break;
        size_t hi = lo + chunk < B->n_reads ? lo + chunk : B->n_reads;
        for (size_t r = lo; r < hi; r++) {
            float fr;
            uint8_t fl;
            long n;
            for (;;) {
                n = orc_map_read(B->idx, B->seqs + B->offsets[r], B->lens[r], B->rescue_cutoff,
                                 B->hard_cutoff, buf, cap, &fr, &fl);
                if (n >= 0) break;
                cap *= 4;
                buf = (orc_nam *)realloc(buf, cap * sizeof(orc_nam));
            }
            if (B->nam_counts) B->nam_counts[r] = (uint32_t)n;
            if (B->fracs) B->fracs[r] = fr;
            if (B->flags) B->flags[r] = fl;
            if (B->nams && B->nam_offsets && n > 0)
                memcpy(B->nams + B->nam_offsets[r], buf, (size_t)n * sizeof(orc_nam));
        }
    }
    free(buf);
    return NULL;
}

int orc_map_batch(const orc_index *idx, const char *seqs, const uint64_t *offsets,
                  const uint32_t *lens, size_t n_reads, unsigned rescue_cutoff,
                  unsigned hard_cutoff, int n_threads, const uint64_t *nam_offsets,
                  orc_nam *nams, uint32_t *nam_counts, float *fracs, uint8_t *flags) {
    batch_job B;
    B.idx = idx; B.seqs = seqs; B.offsets = offsets; B.lens = lens; B.n_reads = n_reads;
    B.rescue_cutoff = rescue_cutoff; B.hard_cutoff = hard_cutoff;
    B.nam_offsets = nam_offsets; B.nams = nams; B.nam_counts = nam_counts;
    B.fracs = fracs; B.flags = flags;
    atomic_store(&B.next, 0);
    atomic_store(&B.err, 0);
    if (n_threads < 1) n_threads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)n_threads);
    for (int t = 0; t < n_threads; t++) pthread_create(&th[t], NULL, batch_worker, &B);
    for (int t = 0; t < n_threads; t++) pthread_join(th[t], NULL);
    free(th);
    return atomic_load(&B.err);
}

/* Algorithmic bytes of the lookup stage for one read's query randstrobes:
 * 16 (the two bucket bounds) + 16 per entry a linear scan touches until the
 * first hash >= key (binary search: 16 per probe) + for a found seed 8 (the
 * filter probe hash) + for an unfiltered seed 16 per entry of its equal-hash
 * run plus the 8-byte hash that ends the run. */
int orc_lookup_bytes(const orc_index *idx, const orc_qrs *q, size_t nq, uint64_t *bytes_out) {
    uint64_t b = 0;
    for (size_t i = 0; i < nq; i++) {
        uint64_t key = q[i].hash;
        const unsigned top_n = (unsigned)(key >> (64 - idx->bits));
        uint64_t ps = idx->starts[top_n], pe = idx->starts[top_n + 1];
        b += 16;
        uint64_t pos = UINT64_MAX;
        for (uint64_t x = ps; x < pe; x++) {
            b += 16;
            if (idx->randstrobes[x].hash >= key) {
                if (idx->randstrobes[x].hash == key) pos = x;
                break;
            }
        }
        if (pos == UINT64_MAX) continue;
        b += 8;
        if (orc_is_filtered(idx, pos)) continue;
        uint64_t x = pos + 1;
        while (x < idx->n && idx->randstrobes[x].hash == key) { b += 16; x++; }
        b += 8;
    }
    *bytes_out = b;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* NAM consumer of paired reads (the step after find_nams in the GPU driver) */

/* quick_sort_iterative, step2/main.cu:317-359: Lomuto partition around the
 * last element, explicit stack, lower part pushed first. The comparator
 * decides where ties land, so the order of equal keys is this algorithm's. */
typedef int (*orc_less_fn)(const void *a, const void *b, const void *ctx);

static void lomuto_sort(void *base, long n, size_t sz, orc_less_fn less, const void *ctx) {
    if (n < 2) return;
    char *d = (char *)base;
    long *stack = (long *)malloc(sizeof(long) * (size_t)(2 * n));
    char *tmp = (char *)malloc(sz);
    char *pivot = (char *)malloc(sz);
    long top = -1;
    stack[++top] = 0;
    stack[++top] = n - 1;
    while (top >= 0) {
        long high = stack[top--];
        long low = stack[top--];
        memcpy(pivot, d + high * sz, sz);
        long i = low - 1;
        for (long j = low; j < high; j++) {
            if (less(d + j * sz, pivot, ctx)) {
                ++i;
                memcpy(tmp, d + i * sz, sz);
                memcpy(d + i * sz, d + j * sz, sz);
                memcpy(d + j * sz, tmp, sz);
            }
        }
        memcpy(tmp, d + (i + 1) * sz, sz);
        memcpy(d + (i + 1) * sz, d + high * sz, sz);
        memcpy(d + high * sz, tmp, sz);
        long p = i + 1;
        if (p - 1 > low) {
            stack[++top] = low;
            stack[++top] = p - 1;
        }
        if (p + 1 < high) {
            stack[++top] = p + 1;
            stack[++top] = high;
        }
    }
    free(stack);
    free(tmp);
    free(pivot);
}

static int imax(int a, int b) { return a > b ? a : b; }

/* Mate 1: sort_nams_single2's lambda, step2/main.cu:2275-2286 (is_rc, ref_id,
 * n_hits descending, query_end, query_start, ref_end, ref_start; true on
 * equality). Mate 2: Nam::operator<, step2/my_struct.hpp:54-68 (is_rc,
 * ref_id, max(0, ref_start - query_start), then as mate 1; true on equality). */
static int nam_ref_less(const orc_nam *a, const orc_nam *b, int mate) {
    if (a->is_rc != b->is_rc) return a->is_rc < b->is_rc;
    if (a->ref_id != b->ref_id) return a->ref_id < b->ref_id;
    if (mate == 2) {
        const int d1 = imax(0, a->ref_start - a->query_start);
        const int d2 = imax(0, b->ref_start - b->query_start);
        if (d1 != d2) return d1 < d2;
    }
    if (a->n_hits != b->n_hits) return a->n_hits > b->n_hits;
    if (a->query_end != b->query_end) return a->query_end < b->query_end;
    if (a->query_start != b->query_start) return a->query_start < b->query_start;
    if (a->ref_end != b->ref_end) return a->ref_end < b->ref_end;
    if (a->ref_start != b->ref_start) return a->ref_start < b->ref_start;
    return 1;
}

static int nam_less_mate1(const void *a, const void *b, const void *ctx) {
    (void)ctx;
    return nam_ref_less((const orc_nam *)a, (const orc_nam *)b, 1);
}
static int nam_less_mate2(const void *a, const void *b, const void *ctx) {
    (void)ctx;
    return nam_ref_less((const orc_nam *)a, (const orc_nam *)b, 2);
}

/* gpu_sort_nams, step2/main.cu:3449-3472: NAM lists of mate 1 (ids below
 * the read count) by sort_nams_single2, of mate 2 by sort_nams_single. */
void orc_sort_nams_reference(orc_nam *nams, size_t n, int mate) {
    lomuto_sort(nams, (long)n, sizeof(orc_nam), mate == 2 ? nam_less_mate2 : nam_less_mate1, NULL);
}

/* top_dropoff, step2/main.cu:889-898 */
float orc_top_dropoff(const orc_nam *nams, size_t n) {
    const orc_nam *n_max = &nams[0];
    if (n_max->n_hits <= 2) return 1.0f;
    if (n > 1) return (float)nams[1].n_hits / (float)n_max->n_hits;
    return 0.0f;
}

/* float to int as the GPU converts (truncation, saturating, NaN -> 0); in the
 * reference's range the plain C conversion */
static int f2i_sat(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return INT32_MAX;
    if (x <= -2147483648.0f) return INT32_MIN;
    return (int)x;
}

/* get_mapq, step2/main.cu:900-910. log is the correctly rounded natural
 * logarithm in float, taken as log in double rounded to float. The uint8
 * return type wraps a negative value. */
uint8_t orc_get_mapq(const orc_nam *nams, size_t n, const orc_nam *n_max) {
    if (n <= 1) return 60;
    const float s1 = n_max->score;
    const float s2 = nams[1].score;
    const double mm = n_max->n_hits / 10.0;
    const float min_matches = (float)(mm < 1.0 ? mm : 1.0);
    const float lg = (float)log((double)s1);
    const float v = 40.0f * (1.0f - s2 / s1) * min_matches * lg;
    const int uncapped = f2i_sat(v);
    return (uint8_t)(uncapped < 60 ? uncapped : 60);
}

/* is_proper_nam_pair, step2/main.cu:870-887 */
int orc_is_proper_nam_pair(const orc_nam *nam1, const orc_nam *nam2, float mu, float sigma) {
    if (nam1->ref_id != nam2->ref_id || nam1->is_rc == nam2->is_rc) return 0;
    const int a = imax(0, nam1->ref_start - nam1->query_start);
    const int b = imax(0, nam2->ref_start - nam2->query_start);
    /* mu + 10 sigma as one fused multiply-add, as the reference's device
     * compiler contracts it */
    const float lim = fmaf(10.0f, sigma, mu);
    const int r1_r2 = nam2->is_rc && (a <= b) && ((float)(b - a) < lim);
    const int r2_r1 = nam1->is_rc && (b <= a) && ((float)(a - b) < lim);
    return r1_r2 || r2_r1;
}

/* gpu_pre_cal_type's body, step2/main.cu:3565-3578: 0 no NAMs, 1 only mate 1,
 * 2 only mate 2, 3 both dominated by one NAM forming a proper pair, 4 the
 * rest. Lists in the reference order. */
int orc_pair_type(const orc_nam *n1, size_t c1, const orc_nam *n2, size_t c2, float dropoff,
                  float mu, float sigma) {
    if (!c1 && !c2) return 0;
    if (c1 && !c2) return 1;
    if (!c1 && c2) return 2;
    if (orc_top_dropoff(n1, c1) < dropoff && orc_top_dropoff(n2, c2) < dropoff &&
        orc_is_proper_nam_pair(&n1[0], &n2[0], mu, sigma))
        return 3;
    return 4;
}

typedef struct {
    const orc_nam *n1, *n2;
} pair_ctx;

/* the NamPair comparator of get_best_scoring_nam_pairs, step2/main.cu:1790-1828;
 * an absent NAM compares as Nam{} with ref_start = -1 */
static int nam_pair_less(const void *pa, const void *pb, const void *vctx) {
    const orc_nam_pair *A = (const orc_nam_pair *)pa, *B = (const orc_nam_pair *)pb;
    const pair_ctx *c = (const pair_ctx *)vctx;
    if (A->score != B->score) return A->score > B->score;
    orc_nam dummy;
    memset(&dummy, 0, sizeof dummy);
    dummy.ref_start = -1;
    const orc_nam *a1 = A->i1 == -1 ? &dummy : &c->n1[A->i1];
    const orc_nam *a2 = A->i2 == -1 ? &dummy : &c->n2[A->i2];
    const orc_nam *b1 = B->i1 == -1 ? &dummy : &c->n1[B->i1];
    const orc_nam *b2 = B->i2 == -1 ? &dummy : &c->n2[B->i2];
    if (a1->score != b1->score) return a1->score > b1->score;
    if (a1->is_rc != b1->is_rc) return !a1->is_rc;
    if (a1->query_end != b1->query_end) return a1->query_end < b1->query_end;
    if (a1->query_start != b1->query_start) return a1->query_start < b1->query_start;
    if (a1->ref_end != b1->ref_end) return a1->ref_end < b1->ref_end;
    if (a1->ref_start != b1->ref_start) return a1->ref_start < b1->ref_start;
    if (a2->score != b2->score) return a2->score > b2->score;
    if (a2->is_rc != b2->is_rc) return !a2->is_rc;
    if (a2->query_end != b2->query_end) return a2->query_end < b2->query_end;
    if (a2->query_start != b2->query_start) return a2->query_start < b2->query_start;
    if (a2->ref_end != b2->ref_end) return a2->ref_end < b2->ref_end;
    return a2->ref_start < b2->ref_start;
}

/* get_best_scoring_nam_pairs, step2/main.cu:1719-1830, on two non-empty NAM
 * lists. Returns the entry count, or -1 if cap is too small or a list is
 * empty. At most 4 * max_tries + 1 entries. */
long orc_best_scoring_nam_pairs(const orc_nam *n1, size_t c1, const orc_nam *n2, size_t c2, float mu,
                                float sigma, int max_tries, orc_nam_pair *out, size_t cap) {
    if (!c1 || !c2 || max_tries < 1) return -1;
    uint8_t *added1 = (uint8_t *)calloc(c1, 1), *added2 = (uint8_t *)calloc(c2, 1);
    long size = 0;
    int ok = 1;
    int best = 0;
#define PUSH(sc, a, b)                                 \
    do {                                               \
        if ((size_t)size >= cap) { ok = 0; goto done; } \
        out[size].score = (sc);                        \
        out[size].i1 = (a);                            \
        out[size].i2 = (b);                            \
        size++;                                        \
    } while (0)
    for (size_t i = 0; i < c1; i++) {
        for (size_t j = 0; j < c2; j++) {
            const int joint = n1[i].n_hits + n2[j].n_hits;
            if ((double)joint < 0.5 * best || size > (long)max_tries * 2) break;
            if (orc_is_proper_nam_pair(&n1[i], &n2[j], mu, sigma)) {
                PUSH(joint, (int)i, (int)j);
                added1[i] = 1;
                added2[j] = 1;
                if (joint > best) best = joint;
            }
        }
        if (size > (long)max_tries * 2) break;
    }
    {
        const int best1 = best > 0 ? best : n1[0].n_hits;
        const size_t lim1 = c1 < (size_t)max_tries ? c1 : (size_t)max_tries;
        for (size_t i = 0; i < lim1; i++) {
            if (n1[i].n_hits < best1 / 2) break;
            if (added1[i]) continue;
            PUSH(n1[i].n_hits, (int)i, -1);
        }
        const int best2 = best > 0 ? best : n2[0].n_hits;
        const size_t lim2 = c2 < (size_t)max_tries ? c2 : (size_t)max_tries;
        for (size_t i = 0; i < lim2; i++) {
            if (n2[i].n_hits < best2 / 2) break;
            if (added2[i]) continue;
            PUSH(n2[i].n_hits, -1, (int)i);
        }
    }
#undef PUSH
    {
        pair_ctx c = {n1, n2};
        lomuto_sort(out, size, sizeof(orc_nam_pair), nam_pair_less, &c);
    }
done:
    free(added1);
    free(added2);
    return ok ? size : -1;
}

/* The pair step over n_pairs pairs (mate 1 of pair p is list p, mate 2 list
 * n_pairs + p): lists copied to ord in the reference order, then per pair the
 * type (gpu_pre_cal_type), per list top_dropoff (0 for an empty list) and
 * get_mapq(list, list[0]), and with joint != 0 the joint entries of pairs
 * with NAMs on both mates at joint_out + p * (4 max_tries + 1). */
int orc_pairs_batch(const uint64_t *nam_offsets, const orc_nam *nams, uint32_t n_pairs, float dropoff,
                    float mu, float sigma, int max_tries, int joint, orc_nam *ord, uint8_t *types,
                    float *drop, uint8_t *mapq, uint32_t *joint_counts, orc_nam_pair *joint_out) {
    const uint64_t n_lists = 2ull * n_pairs;
    for (uint64_t r = 0; r < n_lists; r++) {
        const uint64_t o = nam_offsets[r], c = nam_offsets[r + 1] - o;
        memcpy(ord + o, nams + o, sizeof(orc_nam) * c);
        orc_sort_nams_reference(ord + o, c, r < n_pairs ? 1 : 2);
        drop[r] = c ? orc_top_dropoff(ord + o, c) : 0.0f;
        mapq[r] = orc_get_mapq(ord + o, c, ord + o);
    }
    const size_t stride = 4 * (size_t)max_tries + 1;
    for (uint32_t p = 0; p < n_pairs; p++) {
        const uint64_t o1 = nam_offsets[p], c1 = nam_offsets[p + 1] - o1;
        const uint64_t o2 = nam_offsets[n_pairs + p], c2 = nam_offsets[n_pairs + p + 1] - o2;
        types[p] = (uint8_t)orc_pair_type(ord + o1, c1, ord + o2, c2, dropoff, mu, sigma);
        if (!joint) continue;
        joint_counts[p] = 0;
        if (c1 && c2) {
            long k = orc_best_scoring_nam_pairs(ord + o1, c1, ord + o2, c2, mu, sigma, max_tries,
                                                joint_out + (size_t)p * stride, stride);
            if (k < 0) return ORC_CAPACITY;
            joint_counts[p] = (uint32_t)k;
        }
    }
    return ORC_OK;
}
