/*
 * gis — MI355X-native strobealign seeding path behind a C ABI.
 *
 * syncmers -> randstrobes -> index lookup -> NAMs, computed by hand-written
 * gfx950 kernels. Every entry point returns a gis_status (0 = GIS_OK) and
 * never throws; gis_last_error() gives the message of the last failure on a
 * context. Buffers are caller-owned unless stated. All citations are
 * file:line in the reference repository (yanlifeng/GPU_index_test).
 *
 * Threading: a gis_ctx serialises its calls with an internal mutex; use one
 * context per host thread for concurrency.
 */
#ifndef GIS_H
#define GIS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GIS_ABI_VERSION 1
#define GIS_DEFAULT ((int32_t)0x80000000) /* IndexParameters::DEFAULT, step2/indexparameters.hpp:78 */
#define GIS_NOT_FOUND UINT64_MAX           /* StrobemerIndex::end(), step2/index.hpp:149-151 */
#define GIS_MAX_READ_LEN 65535

typedef enum {
    GIS_OK = 0,
    GIS_ERR_BAD_PARAMETER = 1, /* BadParameter, step2/exceptions.hpp:7-10 */
    GIS_ERR_INVALID_INDEX = 2, /* InvalidIndexFile, step2/exceptions.hpp:17-20 */
    GIS_ERR_HIP = 3,
    GIS_ERR_NOMEM = 4,
    GIS_ERR_CAPACITY = 5,      /* caller buffer too small; the required size is reported */
    GIS_ERR_STATE = 6,         /* e.g. no index loaded */
    GIS_ERR_IO = 7,
    GIS_ERR_BUDGET = 8,        /* request larger than the serving HBM budget */
    GIS_ERR_CANCELLED = 9,
    GIS_ERR_INVALID_FASTA = 10, /* InvalidFasta, step2/exceptions.hpp:12-15 */
    GIS_ERR_INVALID_INPUT = 11  /* malformed FASTQ */
} gis_status;

/* ---------------- plain data types (layouts match the reference) ---------------- */

/* IndexParameters, step2/indexparameters.hpp:11-95 (flattened) */
typedef struct {
    int32_t canonical_read_length;
    int32_t k, s, t_syncmer;
    int32_t l, u;
    uint64_t q;
    int32_t max_dist;
    uint32_t w_min, w_max;
    int32_t reserved_;
} gis_params;

/* Syncmer, step2/randstrobes.hpp:125-131 */
typedef struct { uint64_t hash; uint64_t position; } gis_syncmer;

/* QueryRandstrobe, step2/randstrobes.hpp:96-101 (24 bytes) */
typedef struct {
    uint64_t hash;
    uint32_t start;
    uint32_t end;
    uint8_t is_reverse;
    uint8_t pad_[7];
} gis_query_randstrobe;

/* RefRandstrobe, step2/randstrobes.hpp:65-94: packed = ref_index << 8 | strobe2 offset */
typedef struct { uint64_t hash; uint32_t position; uint32_t packed; } gis_ref_randstrobe;

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
} gis_nam;

/* Hit, step2/my_struct.hpp:71-86, plus the reference index and strand it is
 * grouped under (hits_per_ref[is_revcomp][ref_id], step2/nam.cu:565) */
typedef struct {
    int32_t query_start, query_end, ref_start, ref_end;
    int32_t ref_id;
    int32_t is_rc;
} gis_hit;

typedef struct {
    uint64_t n_randstrobes;
    int32_t bits;
    uint32_t filter_cutoff;
    gis_params params;
    uint64_t distinct_strobemers; /* IndexCreationStatistics, step2/index.hpp:22-35 */
    uint64_t tot_occur_once, tot_high_ab, tot_mid_ab, index_cutoff;
    uint64_t device_bytes;
} gis_index_info;

/* Per-read result flags. GIS_FLAG_SPILLED: in GIS_MODE_FUSED the read did not
 * fit the fused kernel's LDS budget and went through the per-stage kernels
 * (the NAMs are the same either way). */
enum { GIS_FLAG_RESCUED = 1, GIS_FLAG_NO_RANDSTROBES = 2, GIS_FLAG_SPILLED = 0x80 };

/* Output of the NAM entry points. The caller provides arrays; nam_offsets has
 * n_reads + 1 entries and the NAMs of read r are nams[nam_offsets[r] ..
 * nam_offsets[r+1]) in canonical order (is_rc, ref_id, query_start,
 * ref_start) with nam_id = rank. If nam_capacity is too small the call
 * returns GIS_ERR_CAPACITY and sets n_nams to the required count. */
typedef struct {
    uint64_t *nam_offsets;
    gis_nam *nams;
    uint64_t nam_capacity;
    float *nonrepetitive_fraction; /* optional, per read */
    uint8_t *flags;                /* optional, per read */
    uint64_t n_nams;               /* out */
} gis_nam_out;

/* Options of the mapping step. rescue_level is -R (step2/cmdline.hpp:56);
 * rescue_cutoff = R < 100 ? R * filter_cutoff : 100 (step2/main.cu:4204);
 * hard_cutoff is the count limit of find_nams_rescue (step2/nam.cu:964). */
typedef struct {
    int32_t rescue_level;  /* default 2 */
    uint32_t hard_cutoff;  /* default 1000 */
    int32_t mode;          /* GIS_MODE_* */
    int32_t flags;         /* GIS_RUN_* */
} gis_map_options;

/* GIS_MODE_PLAIN: one launch per stage (seed, lookup, hits, rescue, merge).
 * GIS_MODE_FUSED: one wave-per-read kernel for the whole path; reads beyond
 * its LDS budget go through the per-stage kernels. Same NAMs either way. */
enum { GIS_MODE_PLAIN = 0, GIS_MODE_FUSED = 1 };
/* GIS_RUN_NO_GRAPH: enqueue the launches one by one instead of replaying the
 * batch's captured HIP graph. */
enum { GIS_RUN_NO_GRAPH = 1 };

typedef struct gis_ctx gis_ctx;

/* Launch configuration of the fused form (GIS_MODE_FUSED), kept per context.
 * The fused kernel is persistent over a device work queue of read ids sorted
 * by length, longest first; a wave merges the reads it takes four at a time,
 * so persistent grids (blocks_per_cu > 0) suit it. A new context starts with
 * {4, 1, 0, 0, 8}. */
typedef struct {
    int32_t waves_per_block; /* 2, 4 or 8 reads in flight per workgroup */
    int32_t length_split;    /* 1: reads of at most 160 bp run in a second launch with 160-base LDS tiles */
    int32_t queue;           /* 0: static round robin over the queue; 1: dynamic grabs from a device counter */
    int32_t chunk;           /* queue entries per grab when queue = 1 (1..4096) */
    int32_t blocks_per_cu;   /* 0: one wave per read; 1..64: a persistent grid of that many blocks per CU */
} gis_fused_config;

/* ---------------- fused-form configuration (no reference counterpart) ---------------- */
int gis_fused_config_get(gis_ctx *ctx, gis_fused_config *out);
int gis_fused_config_set(gis_ctx *ctx, const gis_fused_config *cfg);

/* ---------------- parameters ---------------- */

/* IndexParameters::from_read_length, step2/indexparameters.hpp:87-89,
 * step2/indexparameters.cpp:51-83. Pass GIS_DEFAULT for defaults. */
int gis_params_from_read_length(int read_length, int k, int s, int l, int u, int c,
                                int max_seed_len, gis_params *out);
/* IndexParameters(r, k, s, l, u, q, max_dist), step2/indexparameters.hpp:80-85 */
int gis_params_make(int canonical_read_length, int k, int s, int l, int u, int q,
                    int max_dist, gis_params *out);
void gis_map_options_default(gis_map_options *opt);

/* ---------------- context ---------------- */
int gis_ctx_create(int device, gis_ctx **out);
void gis_ctx_destroy(gis_ctx *ctx);
const char *gis_last_error(const gis_ctx *ctx);
int gis_device_count(int *count);

/* ---------------- index (StrobemerIndex, step2/index.hpp:37-183) ---------------- */

/* StrobemerIndex::populate(f, n_threads), step2/index.hpp:54 (body
 * step2/index.cu:143-311). bits = -1 picks bits from the total length
 * (step2/index.cu:137-141). The contigs are copied to HBM and seeded there
 * (gis_reference_randstrobes); the records are sorted, bucketed and counted
 * on the GPU, and the index is left resident in HBM. n_threads is the
 * reference's argument and is not needed: no host thread seeds. Each contig
 * must be shorter than 2^32 bases, and there are fewer than 2^24 of them. */
int gis_index_build(gis_ctx *ctx, const char *const *seqs, const uint64_t *lens,
                    uint32_t n_refs, const gis_params *params, int bits, float f,
                    int n_threads);
/* The unsorted RefRandstrobe records of a reference in (contig, position)
 * order: the reference's RandstrobeGenerator over every contig
 * (step2/randstrobes.cu:161-190, count/assign_randstrobes at
 * step2/index.cu:246-311), seeded on the context's GPU. A contig shorter
 * than w_max gets zero records in its slots (the reference counts them but
 * never assigns them). n_out receives the record count; GIS_ERR_CAPACITY if
 * it exceeds capacity. */
int gis_reference_randstrobes(gis_ctx *ctx, const char *const *seqs, const uint64_t *lens, uint32_t n_refs,
                              const gis_params *params, gis_ref_randstrobe *out, uint64_t capacity,
                              uint64_t *n_out);
/* Upload a prebuilt table (e.g. from another builder). The table must be
 * sorted by (hash, position) and starts must have 2^bits + 1 entries. */
int gis_index_upload(gis_ctx *ctx, const gis_ref_randstrobe *randstrobes, uint64_t n,
                     const uint64_t *starts, int bits, uint32_t filter_cutoff,
                     const gis_params *params);
/* StrobemerIndex::read / write, .sti v2 format, step2/index.cu:73-134. When
 * expected is non-NULL the parameters in the file must equal it. */
int gis_index_read_sti(gis_ctx *ctx, const char *path, const gis_params *expected);
int gis_index_write_sti(gis_ctx *ctx, const char *path);
int gis_index_info_get(gis_ctx *ctx, gis_index_info *out);
/* Copy the resident table back (for checks). Either pointer may be NULL. */
int gis_index_download(gis_ctx *ctx, gis_ref_randstrobe *randstrobes, uint64_t *starts);

/* ---------------- stage entry points (host buffers) ---------------- */
/* Reads are concatenated in seqs; read r is seqs[offsets[r] .. offsets[r] + lens[r]). */

/* canonical syncmers (SyncmerIterator, step2/randstrobes.cu:18-79) */
int gis_syncmers_batch(gis_ctx *ctx, const char *seqs, const uint64_t *offsets,
                       const uint32_t *lens, uint32_t n_reads, uint64_t *out_offsets,
                       gis_syncmer *out, uint64_t capacity, uint64_t *n_out);
/* randstrobes_query, step2/randstrobes.hpp:107 (body step2/randstrobes.cu:195-241) */
int gis_randstrobes_batch(gis_ctx *ctx, const char *seqs, const uint64_t *offsets,
                          const uint32_t *lens, uint32_t n_reads, uint64_t *out_offsets,
                          gis_query_randstrobe *out, uint64_t capacity, uint64_t *n_out);
/* gis_lookup_batch without allocating: no reference counterpart. For callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). The copies of the hashes
 * and positions go through one static device array of 2 x
 * GIS_NOALLOC_LOOKUP_MAX u64 (16 MiB) that each GPU holds until the process
 * ends; more than GIS_NOALLOC_LOOKUP_MAX hashes is GIS_ERR_CAPACITY. Calls
 * on one context are serialised; calls on two contexts of the same GPU at
 * the same time race on the array. */
#define GIS_NOALLOC_LOOKUP_MAX (1u << 20)
int gis_lookup_batch_noalloc(gis_ctx *ctx, const uint64_t *hashes, uint64_t n, uint64_t *positions);
/* randstrobes_query(seq, IndexParameters), step2/randstrobes.hpp:107, with
 * the given parameters and no index (a context with or without one). The
 * parameters must be as gis_params_make or gis_params_from_read_length make
 * them (GIS_ERR_BAD_PARAMETER otherwise). */
int gis_randstrobes_params(gis_ctx *ctx, const gis_params *params, const char *seqs, const uint64_t *offsets,
                           const uint32_t *lens, uint32_t n_reads, uint64_t *out_offsets,
                           gis_query_randstrobe *out, uint64_t capacity, uint64_t *n_out);
/* StrobemerIndex::find over a list of hashes, step2/index.hpp:57-81 (the
 * microbenchmark kernel gpu_find, gpu_test.cu:186-215). positions[i] is
 * GIS_NOT_FOUND when absent. */
int gis_lookup_batch(gis_ctx *ctx, const uint64_t *hashes, uint64_t n, uint64_t *positions);
/* Hits of each read before NAM merging (hits_per_ref of find_nams,
 * step2/nam.cu:565-708, or of the rescue path when the read is rescued),
 * grouped by (is_rc, ref_id) and ordered by (query_start, ref_start). */
int gis_hits_batch(gis_ctx *ctx, const char *seqs, const uint64_t *offsets,
                   const uint32_t *lens, uint32_t n_reads, const gis_map_options *opt,
                   uint64_t *out_offsets, gis_hit *out, uint64_t capacity, uint64_t *n_out);
/* find_nams (+ the caller's rescue decision and find_nams_rescue) for every
 * read: step2/nam.hpp:11-20, step2/main.cu:3209-3216. */
int gis_find_nams_batch(gis_ctx *ctx, const char *seqs, const uint64_t *offsets,
                        const uint32_t *lens, uint32_t n_reads, const gis_map_options *opt,
                        gis_nam_out *out);
/* Per-query forms over given QueryRandstrobes (one read's worth or many):
 * query_offsets[n_queries + 1] partitions qrs into reads. The reference
 * calls find_nams once per read from many threads (step2/nam.cu:553-562);
 * calls made at the same time on one context are combined: the first caller
 * to find the context free runs every waiting call with the same rescue
 * settings as one batch and each caller gets its own reads' results. A
 * query's end must be at most GIS_MAX_READ_LEN and its start at most its
 * end; otherwise GIS_ERR_BAD_PARAMETER. */
int gis_find_nams_qrs(gis_ctx *ctx, const gis_query_randstrobe *qrs,
                      const uint64_t *query_offsets, uint32_t n_queries, int rescue,
                      uint32_t rescue_cutoff, uint32_t hard_cutoff, gis_nam_out *out);

/* ---------------- device-resident batches (inputs already in HBM) ---------------- */
typedef struct gis_batch gis_batch;
/* Capacity for up to max_reads reads with max_bases bases in total. */
int gis_batch_create(gis_ctx *ctx, uint32_t max_reads, uint64_t max_bases, gis_batch **out);
/* A batch that runs on the same device workspace (intermediates and results)
 * as share_with, and on its stream: for many resident batches run one after
 * another, the workspace is sized once instead of per batch. Batches sharing
 * a workspace are run one at a time and a run's results are downloaded
 * before another of them runs (GIS_ERR_STATE otherwise). */
int gis_batch_create_shared(gis_ctx *ctx, uint32_t max_reads, uint64_t max_bases, gis_batch *share_with,
                            gis_batch **out);
void gis_batch_destroy(gis_batch *b);
int gis_batch_upload(gis_batch *b, const char *seqs, const uint64_t *offsets,
                     const uint32_t *lens, uint32_t n_reads);
/* Run the whole path on the uploaded reads; stays on the device. */
int gis_batch_run(gis_batch *b, const gis_map_options *opt);
/* Overlapped batches. The reference's GPU driver maps one batch after another
 * with a device-wide synchronisation after each stage
 * (step2/main.cu:4396-4729); here a run can be enqueued and completed apart,
 * so that runs of batches on different streams overlap on the device.
 * gis_batch_create_async makes a batch with its own stream and workspace
 * (gis_batch_create_shared batches of it share both). gis_batch_launch
 * enqueues the run gis_batch_run would do and returns; gis_batch_wait
 * completes it (regrowth reruns included) and leaves the batch as
 * gis_batch_run would. Until then the workspace takes no other run, upload,
 * autotune or stage-time query (GIS_ERR_STATE). One host thread drives a
 * context's launches and waits (calls on a context are serialised). */
int gis_batch_create_async(gis_ctx *ctx, uint32_t max_reads, uint64_t max_bases, gis_batch **out);
int gis_batch_launch(gis_batch *b, const gis_map_options *opt);
int gis_batch_wait(gis_batch *b);
/* Copy results out (canonical NAM order). */
int gis_batch_download(gis_batch *b, gis_nam_out *out);
/* Per-stage kernel times (ms) of the last run, measured with HIP events on the
 * batch's stream. names is filled with a comma-separated list. */
int gis_batch_stage_times(gis_batch *b, double *ms, int cap, int *n, char *names, int names_cap);
/* Algorithmic table bytes of the lookup stage in the last GIS_MODE_PLAIN
 * run, in the byte model of the reference's search (StrobemerIndex::find,
 * is_filtered and the run walk of add_to_hits_per_ref; step2/index.hpp:57-93,
 * step2/nam.cu:45-62): per randstrobe 16 (bucket bounds) + 16 per record a
 * linear scan reads up to the first hash >= key; a found seed 8 more (the
 * filter probe's hash), a kept seed 16 per record of its equal-hash run past
 * the first plus 8 (the hash that ends the run). n_lookups = randstrobes. */
int gis_batch_lookup_bytes(gis_batch *b, uint64_t *bytes, uint64_t *n_lookups);
/* Bytes the lookup kernel of that run streamed beside the table: 12 per
 * query randstrobe read (hash, packed start and end), 12 per result written
 * (position, hit count; only for reads a later stage reads them for: rescued
 * reads, reads whose hits go to the pool, reads of more than 64
 * randstrobes), 16 per hit written into its read's hit region. */
int gis_batch_lookup_streams(gis_batch *b, uint64_t *stream_bytes);
/* total NAMs of the last run (available after gis_batch_run, before download) */
int gis_batch_nam_count(gis_batch *b, uint64_t *n_nams);
/* Autotuner of the fused form: runs the uploaded reads under every candidate
 * gis_fused_config (eager launches, best of three wall-clock runs each),
 * keeps the fastest in the context and reports it. Results are the same
 * under every configuration. */
int gis_batch_autotune(gis_batch *b, const gis_map_options *opt, gis_fused_config *best,
                       double *best_ms, uint32_t *n_tried);

/* ---------------- host-to-host mapping of many reads ---------------- */
/* The reference GPU driver's per-batch loop (step2/main.cu:4396-4729: pack a
 * batch of reads, copy it to the device, run seeding through NAMs, hand the
 * NAMs to the host; rescue as in gis_find_nams_batch) over n_reads reads in
 * host memory, cut in order into batches of batch_reads. Up to `lanes`
 * batches are in flight, each on its own stream and workspace: one batch's
 * copies overlap another's kernels. Results land in out as for
 * gis_find_nams_batch, over all n_reads reads (nam_offsets has n_reads + 1
 * entries). GIS_ERR_CAPACITY (n_nams = the required count) if nam_capacity
 * is too small; every batch still runs. The copies run on two SDMA engines
 * of the GPU, one per direction, so that one batch's reads going in and
 * another's NAMs coming out cross PCIe at the same time; host memory from
 * gis_host_alloc (pinned) is what the engines reach, other memory is copied
 * through HIP, slower. batch_ms (optional, one entry per batch): the batch's
 * latency on the host's clock, from the issue of its first copy to HBM to the
 * completion of its last copy back. Lanes and their buffers stay with the
 * context for later calls. */
int gis_find_nams_stream(gis_ctx *ctx, const char *seqs, const uint64_t *offsets, const uint32_t *lens,
                         uint64_t n_reads, uint32_t batch_reads, uint32_t lanes, const gis_map_options *opt,
                         gis_nam_out *out, double *batch_ms);
/* How gis_find_nams_stream's copies ran on this context: the engines used
 * (hsa_amd_sdma_engine_id_t bits; 0 when copies go through HIP), copies run
 * on them, copies run through HIP, and copies an engine refused (after which
 * the context copies through HIP). No reference counterpart. */
typedef struct gis_copy_stats {
    uint32_t h2d_engine, d2h_engine;
    uint64_t engine_copies, hip_copies, refused;
} gis_copy_stats;
int gis_stream_copy_stats(gis_ctx *ctx, gis_copy_stats *stats);
/* Pinned host memory (hipHostMalloc) for the arrays of gis_find_nams_stream
 * and the other host-buffer entry points; freed with gis_host_free. */
int gis_host_alloc(uint64_t bytes, void **p);
void gis_host_free(void *p);

/* ---------------- serving (no reference counterpart) ---------------- */
/* A request loop on one context: requests (sets of reads of any number) are
 * queued by the caller; one thread per slot admits them while the device
 * bytes charged to the slots stay within hbm_budget, coalesces the queued
 * requests into one launch sequence (a step, replayed from the slot's
 * captured hipGraph), keeps up to `slots` steps in flight on separate
 * streams, and completes each request with its per-read NAMs. A running
 * step is charged the larger of its estimate and what its slot's arrays
 * held before it; an idle slot what its arrays still hold (up to
 * hbm_budget / slots; idle slots free theirs when a step would not fit
 * otherwise). A request of more than max_step_reads reads, or larger than
 * the budget, is cut into parts that run as steps of their own and are
 * joined again for gis_server_wait.
 * Requests arriving while steps run join the next step on a free slot. A
 * request's NAMs are the same alone, batched with others and cut into
 * parts. The context's index must stay loaded while the server exists. */
typedef struct {
    uint32_t slots;            /* steps in flight, each on its own stream (1..16) */
    uint32_t max_step_reads;   /* reads coalesced into one step (a larger request is cut into steps) */
    uint64_t hbm_budget;       /* device bytes charged to the slots (0: 32 GiB) */
    gis_map_options options;   /* mode, rescue level, hard cutoff; graph replays unless GIS_RUN_NO_GRAPH */
} gis_server_config;

typedef struct {
    uint64_t requests, reads, steps;
    double busy_s;             /* time with at least one step in flight */
    uint64_t peak_bytes;       /* largest device bytes charged to the slots */
} gis_server_stats;

typedef struct gis_server gis_server;
/* cfg may be NULL: 6 slots, 131,072 reads per step, 32 GiB, default options */
int gis_server_create(gis_ctx *ctx, const gis_server_config *cfg, gis_server **out);
/* Stops the workers; requests still queued are cancelled (their waiters get
 * GIS_ERR_CANCELLED). Callers blocked in gis_server_wait return before the
 * server is freed; no other call on the server may be in progress or
 * follow. */
void gis_server_destroy(gis_server *s);
/* Queue a request (the reads are copied) and return its id. GIS_ERR_BUDGET
 * only if a single read's step would exceed the HBM budget. */
int gis_server_submit(gis_server *s, const char *seqs, const uint64_t *offsets, const uint32_t *lens,
                      uint32_t n_reads, uint64_t *id);
/* Wait for request id (timeout_ms < 0: no limit; GIS_ERR_STATE on timeout,
 * or at once when another caller is waiting for the same request), copy its
 * results (out may be NULL; GIS_ERR_CAPACITY with n_nams set if
 * nam_capacity is too small, the request stays) and release it. latency_ms:
 * submission to completion of its last step. */
int gis_server_wait(gis_server *s, uint64_t id, int timeout_ms, gis_nam_out *out, double *latency_ms);
/* Replay a timed trace: request i (reads first[i] .. first[i] + count[i] - 1
 * of the given pool) is submitted arrival_s[i] seconds after the start by a
 * client thread; returns when all are done, with each request's latency
 * (ms, -1 if it failed) and the wall time. */
int gis_server_replay(gis_server *s, const char *seqs, const uint64_t *offsets, const uint32_t *lens,
                      uint32_t n_requests, const double *arrival_s, const uint32_t *first,
                      const uint32_t *count, double *latency_ms, double *wall_s);
int gis_server_stats_get(gis_server *s, gis_server_stats *out);

/* ---------------- NAM consumer of paired reads ----------------
 * The step after the seeding path in the reference's GPU driver: each NAM
 * list in the order gpu_sort_nams leaves it (step2/main.cu:3449-3472), the
 * pair types of gpu_pre_cal_type (step2/main.cu:3552-3581) with top_dropoff
 * and is_proper_nam_pair (step2/main.cu:870-898), the per-type pair lists the
 * host builds from them (step2/main.cu:4641-4646), get_mapq
 * (step2/main.cu:900-910) and get_best_scoring_nam_pairs
 * (step2/main.cu:1719-1830). Pairs follow the reference's batch layout: of
 * 2 n NAM lists, mate 1 of pair p is list p and mate 2 is list n + p. */

/* dropoff_threshold and max_tries: MappingParameters (step2/aln.hpp:61-63);
 * mu, sigma: the estimate a fresh GPUInsertSizeDistribution starts from
 * (step2/main.cu:915-920), which gpu_pre_cal_type uses for every pair. */
typedef struct {
    float dropoff_threshold; /* 0.5 */
    int32_t max_tries;       /* 20; 1 .. 256 */
    float mu, sigma;         /* 300, 100 */
    uint32_t flags;          /* GIS_PAIRS_* */
} gis_pair_params;
/* GIS_PAIRS_JOINT: also get_best_scoring_nam_pairs for every pair with NAMs
 * on both mates (types 3 and 4). */
enum { GIS_PAIRS_JOINT = 1 };
void gis_pair_params_default(gis_pair_params *p);

/* NamPair, step2/main.cu:957-963: the joint hit count and the positions of
 * the two NAMs in the mates' reference-ordered lists (-1: no NAM). */
typedef struct { int32_t score; int32_t i1; int32_t i2; } gis_nam_pair;

/* Output of the pair step; every array is optional. NAM lists are in the
 * reference order (mate 1 lists by is_rc, ref_id, n_hits descending,
 * query_end, query_start, ref_end, ref_start; mate 2 lists by is_rc, ref_id,
 * max(0, ref_start - query_start), then the same), with the nam_id each NAM
 * had on input. type_ids lists the pair ids of type 0, then type 1, ...,
 * ascending within a type. top_dropoff is 0 for an empty list. mapq is
 * get_mapq(list, list[0]), the value the reference's type-3 path reports. */
typedef struct {
    uint64_t *nam_offsets;   /* [2 n + 1] */
    gis_nam *nams;
    uint64_t nam_capacity;
    uint8_t *types;          /* [n] 0 .. 4 */
    uint32_t *type_ids;      /* [n] */
    uint32_t type_counts[5]; /* out */
    float *top_dropoff;      /* [2 n] */
    uint8_t *mapq;           /* [2 n] */
    uint64_t *joint_offsets; /* [n + 1] (all 0 without GIS_PAIRS_JOINT) */
    gis_nam_pair *joint;     /* entries of pair p at joint_offsets[p] .. joint_offsets[p+1] */
    uint64_t joint_capacity;
    uint64_t n_nams, n_joint; /* out */
} gis_pair_out;

/* Run the pair step on the NAMs of the batch's last run (2 n reads, mates laid
 * out as above); results stay on the device until downloaded. */
int gis_batch_pairs(gis_batch *b, const gis_pair_params *pp);
/* Copy the pair step's results out. GIS_ERR_CAPACITY (with n_nams / n_joint
 * set) if nam_capacity or joint_capacity is too small. */
int gis_batch_pairs_download(gis_batch *b, gis_pair_out *out);
/* The pair step on NAM lists from host memory (any order within a list;
 * NAMs of one list must differ in (is_rc, ref_id, query_start, ref_start)). */
int gis_pairs_from_nams(gis_ctx *ctx, const uint64_t *nam_offsets, const gis_nam *nams, uint32_t n_pairs,
                        const gis_pair_params *pp, gis_pair_out *out);

/* ---------------- FASTQ ingestion ----------------
 * The chunked producer/consumer in front of the path (producer_pe_fastq_task
 * and the chunk formatting of step2/main.cu:65-84,4212-4240; chunk reading of
 * step2/FastxStream.h:305-364). A reader thread reads plain or gzip FASTQ in
 * chunks of about chunk_bytes cut after complete records (two files: the same
 * record count from each), parser threads format chunks in parallel, and
 * chunks come out in file order.
 * Records: lines end at '\n' (a '\r' stays in the line, as the reference
 * reads CRLF files), the read is the second line, the name the first line
 * without the leading '@' up to the first space (step2/main.cu:4121-4138).
 * GIS_ERR_INVALID_INPUT for a record not starting with '@', a third line not
 * starting with '+', an incomplete last record, mates files of different
 * record counts, or a read longer than GIS_MAX_READ_LEN. */
typedef struct gis_fastq gis_fastq;

/* A chunk of records. With two files, mate 1 of record i is read i and mate
 * 2 is read n_records + i (the layout of the pair step). Valid until the next
 * gis_fastq_next or gis_fastq_close. */
typedef struct {
    uint32_t n_records;           /* 0: end of input */
    uint32_t n_reads;             /* n_records, or 2 n_records with two files */
    const char *seqs;
    const uint64_t *offsets;      /* [n_reads] */
    const uint32_t *lens;         /* [n_reads] */
    const char *names;
    const uint64_t *name_offsets; /* [n_reads + 1] */
    uint64_t index;               /* chunk number from 0 */
    uint64_t first_record;        /* records before this chunk */
    uint64_t bases;
} gis_fastq_chunk;

/* path2 NULL: single-end. chunk_bytes 0: 4 MiB (the reference's 1 << 22
 * pool buffers, step2/main.cu:4214). threads 0: up to 8 parser threads.
 * GIS_ERR_IO if a file cannot be opened. */
int gis_fastq_open(const char *path1, const char *path2, uint64_t chunk_bytes, uint32_t threads,
                   gis_fastq **out);
int gis_fastq_next(gis_fastq *f, gis_fastq_chunk *out);
const char *gis_fastq_last_error(const gis_fastq *f);
void gis_fastq_close(gis_fastq *f);
/* The next chunk of f submitted to the server as one request (its reads as
 * gis_fastq_next lays them out); id 0 and chunk->n_records 0 at the end. */
int gis_server_submit_fastq(gis_server *s, gis_fastq *f, uint64_t *id, gis_fastq_chunk *chunk);

/* ---------------- synthetic inputs (deterministic) ---------------- */
typedef struct {
    uint64_t total_length;
    uint32_t n_contigs;
    double gc;
    double repeat_fraction;
    uint32_t n_families;
    uint32_t family_min_len, family_max_len;
    uint32_t copies_min, copies_max;
    double div_min, div_max;
    double tandem_fraction;
    uint32_t tandem_unit_min, tandem_unit_max;
    uint32_t contig_min_len, contig_max_len; /* 0: equal split */
    uint64_t seed;
} gis_genome_spec;

typedef struct {
    uint64_t n_reads;           /* reads (single) or pairs (paired) */
    int32_t paired;
    uint32_t lengths[8];        /* length mixture */
    double weights[8];
    uint32_t n_lengths;
    double sub_rate, indel_rate;
    double rc_fraction;         /* single-end: fraction reverse-complemented */
    double insert_mean, insert_sd;
    uint64_t seed;
} gis_read_spec;

/* Genome: concatenated contigs in seq (capacity total_length), contig lengths
 * in contig_lens[n_contigs]. */
int gis_synth_genome(const gis_genome_spec *spec, char *seq, uint64_t *contig_lens);
/* Reads sampled from the genome. Outputs for n_out = n_reads (x2 if paired)
 * reads: seqs (capacity n_out * max length), offsets, lens, and the truth
 * (contig, position, is_rc) of each read. Mates of a pair are adjacent. */
int gis_synth_reads(const gis_read_spec *spec, const char *genome, const uint64_t *contig_lens,
                    uint32_t n_contigs, char *seqs, uint64_t *offsets, uint32_t *lens,
                    uint32_t *truth_contig, uint32_t *truth_pos, uint8_t *truth_rc);

#ifdef __cplusplus
}
#endif
#endif
