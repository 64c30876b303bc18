/*
 * CPU oracle for strobealign's seeding path (syncmers -> randstrobes ->
 * index lookup -> NAMs).
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker that the HIP product is
 * compared against, and the timed CPU baseline. Only tests/, the smoke()
 * entry and bench.py's CPU-baseline leg may load it; the product library
 * never links or calls it.
 *
 * It is a plain-C restatement of the reference's CPU algorithm. Every
 * function cites the reference file:line it follows (paths relative to the
 * reference repository root). Types here are the oracle's own; their byte
 * layouts match the reference structs so results compare field by field.
 */
#ifndef GIS_ORACLE_H
#define GIS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* IndexParameters, step2/indexparameters.hpp:11-95 (flattened). */
typedef struct {
    int32_t canonical_read_length;
    int32_t k, s, t_syncmer;            /* SyncmerParameters */
    int32_t l, u;                       /* RandstrobeParameters */
    uint64_t q;
    int32_t max_dist;
    uint32_t w_min, w_max;
} orc_params;

/* Syncmer, step2/randstrobes.hpp:125-131 */
typedef struct { uint64_t hash; uint64_t position; } orc_syncmer;

/* QueryRandstrobe, step2/randstrobes.hpp:96-101 (24 bytes with padding) */
typedef struct {
    uint64_t hash;
    uint32_t start;
    uint32_t end;
    uint8_t is_reverse;
    uint8_t pad_[7];
} orc_qrs;

/* RefRandstrobe, step2/randstrobes.hpp:65-94 (16 bytes) */
typedef struct { uint64_t hash; uint32_t position; uint32_t packed; } orc_ref_randstrobe;

/* Hit, step2/my_struct.hpp:71-86 */
typedef struct { int32_t query_start, query_end, ref_start, ref_end; } orc_hit;

/* Nam, step2/my_struct.hpp:25-69 (44 bytes) */
typedef struct {
    int32_t nam_id;
    int32_t query_start;
    int32_t query_end;
    int32_t query_prev_hit_startpos;
    int32_t ref_start;
    int32_t ref_end;
    int32_t ref_prev_hit_startpos;
    int32_t n_hits;
    int32_t ref_id;
    float score;
    uint8_t is_rc;
    uint8_t pad_[3];
} orc_nam;

/* StrobemerIndex data, step2/index.hpp:37-183 */
typedef struct {
    orc_ref_randstrobe *randstrobes;
    uint64_t n;
    uint64_t *starts;             /* 2^bits + 1 entries */
    int32_t bits;
    uint32_t filter_cutoff;
    orc_params params;
    /* IndexCreationStatistics subset, step2/index.hpp:22-35 */
    uint64_t distinct_strobemers, tot_occur_once, tot_high_ab, tot_mid_ab, index_cutoff;
    int owns;
} orc_index;

/* error codes */
enum { ORC_OK = 0, ORC_BAD_PARAMETER = 1, ORC_CAPACITY = 2, ORC_NOMEM = 3 };

/* ---- hashing / tables ---- */
uint64_t orc_xxh64(uint64_t input);
int orc_nt4(unsigned char c);

/* ---- parameters ---- */
#define ORC_DEFAULT ((int32_t)0x80000000)
int orc_params_from_read_length(int read_length, int k, int s, int l, int u, int c,
                                int max_seed_len, orc_params *out);
int orc_params_make(int canonical_read_length, int k, int s, int l, int u, int q,
                    int max_dist, orc_params *out);

/* ---- seeding ---- */
size_t orc_syncmers(const char *seq, size_t len, const orc_params *p,
                    orc_syncmer *out, size_t cap);
size_t orc_randstrobes_query(const char *seq, size_t len, const orc_params *p,
                             orc_qrs *out, size_t cap);

/* ---- index ---- */
int orc_pick_bits(uint64_t total_length, const orc_params *p);
int orc_index_build(const char *const *seqs, const uint64_t *lens, uint32_t n_refs,
                    const orc_params *p, int bits, float f, int n_threads,
                    orc_index **out);
int orc_index_wrap(orc_ref_randstrobe *rs, uint64_t n, uint64_t *starts, int bits,
                   uint32_t filter_cutoff, const orc_params *p, orc_index **out);
void orc_index_free(orc_index *idx);
uint64_t orc_find(const orc_index *idx, uint64_t key);
int orc_is_filtered(const orc_index *idx, uint64_t position);
uint32_t orc_get_count(const orc_index *idx, uint64_t position);

/* ---- NAMs ---- */
/* nams are written in the reference's own output order (nam_id = append
 * position); orc_canonicalize re-sorts them. Returns count, or -1 if cap is
 * too small. */
long orc_find_nams(const orc_index *idx, const orc_qrs *q, size_t nq,
                   orc_nam *out, size_t cap, float *nonrep_fraction, uint64_t *n_hits);
long orc_find_nams_rescue(const orc_index *idx, const orc_qrs *q, size_t nq,
                          unsigned rescue_cutoff, unsigned hard_cutoff,
                          orc_nam *out, size_t cap);
/* Hit-outer merge (find_nams' form) and NAM-outer merge (rescue's form) on an
 * explicit hit list; exposed so tests can compare the two merges directly.
 * hits are (ref_id, hit) pairs in insertion order. */
long orc_merge_hits(const int32_t *ref_ids, const orc_hit *hits, size_t nh, int k,
                    int is_revcomp, int nam_outer, orc_nam *out, size_t cap);
void orc_canonicalize(orc_nam *nams, size_t n);

/* ---- whole read / batch (CPU baseline) ---- */
enum { ORC_FLAG_RESCUED = 1 };
long orc_map_read(const orc_index *idx, const char *seq, size_t len,
                  unsigned rescue_cutoff, unsigned hard_cutoff,
                  orc_nam *out, size_t cap, float *frac, uint8_t *flags);
/* nam_counts[n_reads]; nams written at nam_offsets supplied by the caller
 * (pass NULL to only count). Threads pull 10 000-read chunks. */
int orc_map_batch(const orc_index *idx, const char *seqs, const uint64_t *offsets,
                  const uint32_t *lens, size_t n_reads, unsigned rescue_cutoff,
                  unsigned hard_cutoff, int n_threads, const uint64_t *nam_offsets,
                  orc_nam *nams, uint32_t *nam_counts, float *fracs, uint8_t *flags);

/* algorithmic byte accounting for a read (lookup-stage model); see bench.py */
int orc_lookup_bytes(const orc_index *idx, const orc_qrs *q, size_t nq,
                     uint64_t *bytes_out);


/* ---- NAM consumer of paired reads ---- */
/* NamPair, step2/main.cu:957-963, with list indices (-1: no NAM) */
typedef struct { int32_t score, i1, i2; } orc_nam_pair;
/* gpu_sort_nams order of one NAM list (mate 1 or 2) */
void orc_sort_nams_reference(orc_nam *nams, size_t n, int mate);
float orc_top_dropoff(const orc_nam *nams, size_t n);
uint8_t orc_get_mapq(const orc_nam *nams, size_t n, const orc_nam *n_max);
int orc_is_proper_nam_pair(const orc_nam *nam1, const orc_nam *nam2, float mu, float sigma);
int orc_pair_type(const orc_nam *n1, size_t c1, const orc_nam *n2, size_t c2, float dropoff,
                  float mu, float sigma);
long orc_best_scoring_nam_pairs(const orc_nam *n1, size_t c1, const orc_nam *n2, size_t c2, float mu,
                                float sigma, int max_tries, orc_nam_pair *out, size_t cap);
int orc_pairs_batch(const uint64_t *nam_offsets, const orc_nam *nams, uint32_t n_pairs, float dropoff,
                    float mu, float sigma, int max_tries, int joint, orc_nam *ord, uint8_t *types,
                    float *drop, uint8_t *mapq, uint32_t *joint_counts, orc_nam_pair *joint_out);

#ifdef __cplusplus
}
#endif
#endif
