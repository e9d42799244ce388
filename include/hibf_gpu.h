/*
 * hibf_gpu.h -- C-ABI of the MI355X-native HIBF query path.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, never
 * throws, and returns an int status (HIBF_GPU_OK == 0).  Pointers named d_*
 * are device pointers (HBM); pointers named h_* are host pointers.  `stream`
 * is a hipStream_t passed as void* (NULL = the default stream).  Query entry
 * points are asynchronous on `stream` unless stated otherwise.
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the seqan/hibf 1.0.0-rc.1 source tree).
 */
#ifndef HIBF_GPU_H
#define HIBF_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum hibf_gpu_status
{
    HIBF_GPU_OK = 0,
    HIBF_GPU_ERR_INVALID_ARGUMENT = 1, /* null handle/pointer, size mismatch          */
    HIBF_GPU_ERR_BIN_COUNT = 2,        /* "The number of bins must be > 0."           */
    HIBF_GPU_ERR_HASH_COUNT = 3,       /* "The number of hash functions must be > 0 and <= 5." */
    HIBF_GPU_ERR_BIN_SIZE = 4,         /* "The size of a bin must be > 0."            */
    HIBF_GPU_ERR_OUT_OF_MEMORY = 5,    /* device allocation failed                    */
    HIBF_GPU_ERR_DEVICE = 6,           /* a HIP runtime call failed                   */
    HIBF_GPU_ERR_CAPACITY = 7,         /* an output buffer is too small               */
    HIBF_GPU_ERR_UNSUPPORTED = 8,      /* geometry outside what the kernels handle    */
    HIBF_GPU_ERR_CONFIG = 9,           /* invalid config (config.cpp:64-130)          */
    HIBF_GPU_ERR_BUDGET = 10,          /* admission over the counter-memory budget    */
    HIBF_GPU_ERR_NO_DEVICE = 11,       /* no HIP device visible                       */
    HIBF_GPU_ERR_IO = 12               /* file cannot be read/written, or is not an archive of the type */
};

/* Text of a status code (static storage). */
const char * hibf_gpu_status_string(int status);
/* Detail of the last error raised on the calling thread (static per thread). */
const char * hibf_gpu_last_error(void);

/* Launch-shape knobs of the query kernels (no counterpart in the reference:
 * the reference has no device tiles).  Process-wide; set by the autotuner
 * (hibf_amd/autotune.py) or by hand.  Keys: "count.wide_kernel" (0/1),
 * "count.wide_unroll" (values in flight per wave: 1,2,4,8),
 * "count.wide_waves" (waves per query block: 1..4), "count.counter_bits"
 * (LDS counter width: 8 or 16), "hibf.wave_kernel" (query per wave 1, per
 * block 0), "hibf.row_per_lane" (descent over rows of <= 4 words: one value
 * and its whole row per lane 1, lanes per value 0), "hibf.sliced_counts"
 * (descent over rows of <= 4 words: bit-sliced per-lane counts 1, one LDS
 * atomic per set bit 0), "hibf.unroll" (descent over rows of <= 4 words with
 * sliced counts: value groups whose rows a wave loads at once, 2 or 4
 * default), "hibf.nt_loads" (0/1), "hibf.stream_kib" (IBFs of at least this
 * many KiB take non-temporal row loads in the count and in the membership
 * descent, except two-hash membership over rows of more than 4 words and
 * state in device memory; 65536 default, 0 none), "hibf.block_waves" (waves
 * per query block of the membership descent: 1, 2, 4 default, 8), "count.nt_stores"
 * (non-temporal stores of the IBF count kernel's counters, of the HIBF
 * count's zero fill and of bulk_contains' rows: 1, default; plain stores 0),
 * "count.wide_nt_loads" (non-temporal row loads of the wide count kernel for
 * IBFs with 2 or 3 hash functions: 1, default; plain loads 0), "contains.unroll" (values whose rows one wave of bulk_contains loads at
 * once: 1, 2, 4 default, 8).  Results never depend on these values.
 * Two keys change what the library does: "build.split_order" (1 default:
 * HIBF builds chunk split user bins in the reference's set order; 0: by value
 * ranges) and "alloc.max_bytes" (largest single device allocation the
 * library makes, 0 default = no limit; larger requests fail with
 * HIBF_GPU_ERR_OUT_OF_MEMORY, for exercising those paths). */
int hibf_gpu_tune_set(const char * key, int64_t value);
int hibf_gpu_tune_get(const char * key, int64_t * value);

/* Device memory helpers so that host code needs no HIP headers.
 * kind: 0 host->device, 1 device->host, 2 device->device (synchronous). */
int hibf_gpu_malloc(void ** d_ptr, uint64_t bytes);
int hibf_gpu_free(void * d_ptr);
int hibf_gpu_memcpy(void * dst, void const * src, uint64_t bytes, int kind);
int hibf_gpu_synchronize(void);

/* Sentinels of ibf_bin_to_user_bin_id, hierarchical_interleaved_bloom_filter.hpp:32-40. */
#define HIBF_GPU_BIN_MERGED 0xFFFFFFFFFFFFFFFFULL
#define HIBF_GPU_BIN_DELETED 0xFFFFFFFFFFFFFFFEULL

/* ======================================================================== */
/* Interleaved Bloom Filter (device resident)                               */
/* ======================================================================== */
typedef struct hibf_gpu_ibf_s * hibf_gpu_ibf;

/* Geometry as interleaved_bloom_filter.hpp:160-170. */
typedef struct hibf_gpu_ibf_geometry
{
    uint64_t bins;           /* bin_count()                              */
    uint64_t technical_bins; /* 64 * bin_words                            */
    uint64_t bin_size;       /* bin_size(): rows                          */
    uint64_t hash_shift;     /* countl_zero(bin_size)                     */
    uint64_t bin_words;      /* ceil(bins / 64)                           */
    uint64_t hash_funs;      /* hash_function_count()                     */
    uint64_t n_words;        /* technical_bins * bin_size / 64            */
} hibf_gpu_ibf_geometry;

/* Replaces interleaved_bloom_filter(bin_count, bin_size, hash_function_count)
 * (interleaved_bloom_filter.hpp:225-228; src/interleaved_bloom_filter.cpp:32-53):
 * allocates a zeroed bit matrix in HBM.  Errors 2/3/4 mirror the reference's
 * three std::logic_error cases in the same order. */
int hibf_gpu_ibf_create(uint64_t bins, uint64_t bin_size, uint32_t hash_funs, hibf_gpu_ibf * out);
int hibf_gpu_ibf_destroy(hibf_gpu_ibf ibf);

/* Replaces bin_count()/bin_size()/bit_size()/hash_function_count()
 * (interleaved_bloom_filter.hpp:400-427). */
int hibf_gpu_ibf_get_geometry(hibf_gpu_ibf ibf, hibf_gpu_ibf_geometry * out);

/* Device address of the bit matrix (n_words uint64 words, row-major:
 * row r = words [r*bin_words, (r+1)*bin_words)); replaces data()
 * (interleaved_bloom_filter.hpp:440-446). */
int hibf_gpu_ibf_device_words(hibf_gpu_ibf ibf, uint64_t ** d_words);

/* Host <-> device copies of the whole bit matrix (synchronous).  The word
 * layout is the reference's bit_vector storage (bit_vector.hpp:868-876). */
int hibf_gpu_ibf_upload(hibf_gpu_ibf ibf, uint64_t const * h_words, uint64_t n_words);
int hibf_gpu_ibf_download(hibf_gpu_ibf ibf, uint64_t * h_words, uint64_t n_words);

/* Replace try_increase_bin_number_to / increase_bin_number_to
 * (interleaved_bloom_filter.hpp:320-344; src/interleaved_bloom_filter.cpp:150-201).
 * try: when the words per row stay the same only bin_count changes and
 * *increased is 1; otherwise nothing changes and *increased is 0.
 * increase: additionally re-strides the bit matrix on the device to
 * ceil(new_bins/64) words per row (new columns zero), growing the occupancy
 * with zeros; new_bins < bin_count returns HIBF_GPU_ERR_INVALID_ARGUMENT with
 * the reference's message.  Device addresses from hibf_gpu_ibf_device_words
 * are invalid afterwards; synchronous. */
int hibf_gpu_ibf_try_increase_bin_number_to(hibf_gpu_ibf ibf, uint64_t new_bins, int * increased);
int hibf_gpu_ibf_increase_bin_number_to(hibf_gpu_ibf ibf, uint64_t new_bins);

/* Occupancy kept with the handle (interleaved_bloom_filter.hpp:451-458):
 * technical_bins counters and the track_occupancy flag.  With the flag set,
 * hibf_gpu_ibf_emplace updates them exactly as the reference's emplace calls
 * in pair order would; clear() resets the cleared bins' counters as
 * src/interleaved_bloom_filter.cpp:145 does.  h_occupancy may be NULL (get:
 * skipped; set: all zero). */
int hibf_gpu_ibf_occupancy(hibf_gpu_ibf ibf, uint64_t * h_occupancy, int * track_occupancy);
int hibf_gpu_ibf_set_occupancy(hibf_gpu_ibf ibf, uint64_t const * h_occupancy, int track_occupancy);

/* Bulk emplace; replaces emplace(value, bin_index) (interleaved_bloom_filter.hpp:259;
 * src/interleaved_bloom_filter.cpp:119-140) called for n (value, bin) pairs.
 * d_bins may be NULL, in which case every value goes to `bin`.  Bits are set
 * with atomicOr, so the bits do not depend on order or concurrency.
 * Bins >= bin_count are rejected only when d_bins == NULL (checked on the
 * host); per-pair bins are the caller's precondition as in the reference
 * (such pairs are skipped).  With track_occupancy set on the handle, each
 * bin's occupancy grows by the number of its pairs that set a bit unset just
 * before them, taking the pairs in order (d_values[0] first), as n calls of
 * the reference's emplace do; that form sorts the n x h bit positions on the
 * device, is synchronous and takes at most 2^31 - 1 positions per call
 * (HIBF_GPU_ERR_UNSUPPORTED beyond). */
int hibf_gpu_ibf_emplace(hibf_gpu_ibf ibf,
                         uint64_t const * d_values,
                         uint64_t const * d_bins,
                         uint64_t bin,
                         uint64_t n,
                         void * stream);

/* Generator-driven bulk emplace (no counterpart in the reference API; it
 * replaces the caller loop that emplaces a generated set per bin, as the
 * reference's IBF config constructor does through input_fn,
 * src/interleaved_bloom_filter.cpp:104-117): technical bin first_bin + b
 * (b < n_bins) receives hibf_gpu_synth_value(seed, first_bin + b, j) for
 * j < values_per_bin. */
int hibf_gpu_ibf_emplace_synthetic(hibf_gpu_ibf ibf,
                                   uint64_t seed,
                                   uint64_t first_bin,
                                   uint64_t n_bins,
                                   uint64_t values_per_bin,
                                   void * stream);

/* The synthetic value stream: fmix64((bin << 36 | i) + seed * 0x9E3779B97F4A7C15),
 * distinct for distinct (bin, i) with i < 2^36 and bin < 2^28. */
uint64_t hibf_gpu_synth_value(uint64_t seed, uint64_t bin, uint64_t i);

/* Replaces clear(bin_index) / clear(range) (interleaved_bloom_filter.hpp:270, 283-297;
 * src/interleaved_bloom_filter.cpp:142-148) for n bins given on the host. */
int hibf_gpu_ibf_clear(hibf_gpu_ibf ibf, uint64_t const * h_bins, uint64_t n, void * stream);

/* Batched containment_agent_type::bulk_contains (src/interleaved_bloom_filter.cpp:204-291):
 * d_out receives n * bin_words words, the AND of the h rows of each value. */
int hibf_gpu_ibf_bulk_contains(hibf_gpu_ibf ibf,
                               uint64_t const * d_values,
                               uint64_t n,
                               uint64_t * d_out,
                               void * stream);

/* Batched counting_agent_type<T>::bulk_count (interleaved_bloom_filter.hpp:652-665)
 * over nq CSR queries: query q is d_values[d_offsets[q] .. d_offsets[q+1]).
 * d_counts receives nq * bin_count counters of counter_bytes (1, 2, 4 or 8)
 * bytes each, wrapping modulo 2^(8*counter_bytes) like counting_vector<T>. */
int hibf_gpu_ibf_bulk_count(hibf_gpu_ibf ibf,
                            uint64_t const * d_values,
                            uint64_t const * d_offsets,
                            uint64_t nq,
                            uint32_t counter_bytes,
                            void * d_counts,
                            void * stream);

/* ======================================================================== */
/* Query agents: per-caller device workspace (one agent per host thread, as  */
/* the reference asks for its agents, interleaved_bloom_filter.hpp:547-551). */
/* ======================================================================== */
typedef struct hibf_gpu_agent_s * hibf_gpu_agent;

/* max_queries_per_launch bounds the per-call workspace (0 = 65536). */
int hibf_gpu_agent_create(uint64_t max_queries_per_launch, hibf_gpu_agent * out);
int hibf_gpu_agent_destroy(hibf_gpu_agent agent);

/* Work counters of the HIBF descent launched through this agent since the
 * last reset: IBF visits (one per query and visited IBF) and row bytes
 * fetched (query length x h x bin_words x 8 per visit).  Synchronous.  No
 * counterpart in the reference (it has no device traffic to report). */
int hibf_gpu_agent_stats(hibf_gpu_agent agent, uint64_t * visits, uint64_t * lookups, uint64_t * row_bytes,
                         int reset);

/* After hibf_gpu_hibf_membership_launch calls: ids found by the last launch
 * and whether any launch since the last prepare dropped ids because its
 * output was full.  Synchronous. */
int hibf_gpu_agent_status(hibf_gpu_agent agent, uint64_t * last_total, uint64_t * dropped);

/* Batched membership_agent_type::membership_for (interleaved_bloom_filter.hpp:744-764):
 * bins with uint16 count >= d_thresholds[q], ascending.  Results of query q
 * are d_ids[d_hit_offsets[q] .. d_hit_offsets[q+1]) (d_hit_offsets has nq+1
 * entries).  Synchronous: returns HIBF_GPU_ERR_CAPACITY and writes the
 * needed total to *h_total when ids_cap is too small. */
int hibf_gpu_ibf_membership_for(hibf_gpu_ibf ibf,
                                hibf_gpu_agent agent,
                                uint64_t const * d_values,
                                uint64_t const * d_offsets,
                                uint64_t nq,
                                uint16_t const * d_thresholds,
                                uint64_t * d_ids,
                                uint64_t ids_cap,
                                uint64_t * d_hit_offsets,
                                uint64_t * h_total,
                                void * stream);

/* ======================================================================== */
/* Hierarchical Interleaved Bloom Filter (device resident)                  */
/* ======================================================================== */
typedef struct hibf_gpu_hibf_s * hibf_gpu_hibf;

/* Upload of a built HIBF: replaces the public members ibf_vector, next_ibf_id
 * and ibf_bin_to_user_bin_id (hierarchical_interleaved_bloom_filter.hpp:145, 209-238).
 *   geoms[i]            bins, bin_size, hash_funs of IBF i (other fields ignored)
 *   h_arena             concatenated bit matrices; IBF i starts at h_word_offsets[i]
 *                       (NULL: allocate zeroed and fill later with hibf_gpu_hibf_emplace)
 *   h_tb_offsets        n_ibfs+1 CSR offsets into the per-technical-bin tables
 *   h_next_ibf          next_ibf_id[i][b]
 *   h_user_bin          ibf_bin_to_user_bin_id[i][b] (HIBF_GPU_BIN_MERGED/DELETED sentinels)
 * IBFs of any number of technical bins are accepted.  When (1 + depth) x the
 * largest IBF's technical bins x 4 bytes exceeds 64 KiB, the descent keeps a
 * query's per-IBF counters and per-level event lists in device memory (that
 * many bytes per block in the agent's workspace) instead of LDS.  Limits:
 * depth <= 16, fewer than 2^31 IBFs and user bins, word offsets multiples of
 * 4 (HIBF_GPU_ERR_UNSUPPORTED / _INVALID_ARGUMENT otherwise). */
int hibf_gpu_hibf_create(uint64_t n_ibfs,
                         hibf_gpu_ibf_geometry const * geoms,
                         uint64_t const * h_arena,
                         uint64_t const * h_word_offsets,
                         uint64_t const * h_tb_offsets,
                         uint64_t const * h_next_ibf,
                         uint64_t const * h_user_bin,
                         uint64_t number_of_user_bins,
                         hibf_gpu_hibf * out);
int hibf_gpu_hibf_destroy(hibf_gpu_hibf hibf);

/* Bulk emplace into IBF `ibf_idx` of the HIBF (same contract as hibf_gpu_ibf_emplace). */
int hibf_gpu_hibf_emplace(hibf_gpu_hibf hibf,
                          uint64_t ibf_idx,
                          uint64_t const * d_values,
                          uint64_t const * d_bins,
                          uint64_t bin,
                          uint64_t n,
                          void * stream);

/* Copies the whole device arena back (synchronous, for parity checks). */
int hibf_gpu_hibf_download_arena(hibf_gpu_hibf hibf, uint64_t * h_arena, uint64_t n_words);
int hibf_gpu_hibf_arena_words(hibf_gpu_hibf hibf, uint64_t * n_words);
/* Device address of the arena (IBF i's words start at word_offsets[i]). */
int hibf_gpu_hibf_device_arena(hibf_gpu_hibf hibf, uint64_t ** d_arena);

/* The layout this HIBF was built from, in the reference's layout text format
 * (src/layout/layout.cpp:125-136): the given one or the one the build
 * computed; empty for an HIBF created from tables or loaded.  No counterpart
 * in the reference's HIBF, which does not keep its layout.  Writes the text
 * and its terminating NUL; HIBF_GPU_ERR_CAPACITY + *needed when cap is too
 * small. */
int hibf_gpu_hibf_layout(hibf_gpu_hibf hibf, char * text, uint64_t cap, uint64_t * needed);

/* Phase times of the build that made this HIBF, in seconds (all 0 for an HIBF
 * created from tables or loaded): [0] user-bin sketches and their estimates,
 * [1] layout, [2] allocation of the arena and tables, [3] filling the IBFs on
 * the device.  They feed the reference's timers
 * (hierarchical_interleaved_bloom_filter.hpp:283-290). */
int hibf_gpu_hibf_build_times(hibf_gpu_hibf hibf, double * h_seconds);

/* Parts of the layout time of that build, in seconds: [0] union estimation
 * (on the device), [1] similarity rearrangement.  Both 0 when the layout did
 * not use union estimates.  They feed layout_union_estimation_timer and
 * layout_rearrangement_timer (hierarchical_interleaved_bloom_filter.hpp:283-290;
 * filled by compute_layout, src/layout/compute_layout.cpp:26-55). */
int hibf_gpu_hibf_layout_times(hibf_gpu_hibf hibf, double * h_seconds);

/* ibf_vector[i] (hierarchical_interleaved_bloom_filter.hpp:209): an IBF handle
 * whose bit matrix is IBF i's words in the HIBF's arena (no copy).  Queries,
 * emplace, clear, upload and download act on the arena; the bin count cannot
 * change (try_/increase_bin_number_to return HIBF_GPU_ERR_UNSUPPORTED).
 * Destroying the view frees only the handle; it must not outlive the HIBF. */
int hibf_gpu_hibf_ibf_view(hibf_gpu_hibf hibf, uint64_t i, hibf_gpu_ibf * out);

/* Batched HIBF membership_agent_type::membership_for
 * (hierarchical_interleaved_bloom_filter.hpp:321-355, 424-439): user-bin ids in
 * the reference's unsorted depth-first order.  Same output contract as
 * hibf_gpu_ibf_membership_for (synchronous; HIBF_GPU_ERR_CAPACITY + *h_total). */
int hibf_gpu_hibf_membership_for(hibf_gpu_hibf hibf,
                                 hibf_gpu_agent agent,
                                 uint64_t const * d_values,
                                 uint64_t const * d_offsets,
                                 uint64_t nq,
                                 uint16_t const * d_thresholds,
                                 uint64_t * d_ids,
                                 uint64_t ids_cap,
                                 uint64_t * d_hit_offsets,
                                 uint64_t * h_total,
                                 void * stream);

/* Asynchronous, graph-capturable form of hibf_gpu_hibf_membership_for: one
 * persistent work-stealing launch over nq queries.  Query q's ids (DFS order)
 * are d_ids[d_hit_start[q] .. d_hit_start[q] + d_hit_count[q]); ids beyond
 * ids_cap are dropped (hibf_gpu_agent_status reports it).  The agent must
 * first be prepared (allocates its workspace; not capturable) for batches of
 * up to max_queries queries, and prepared again after "hibf.block_waves" or
 * "hibf.wave_kernel" change (a launch on a workspace sized for another shape
 * returns HIBF_GPU_ERR_INVALID_ARGUMENT).  Calls on one agent must be
 * serialised. */
int hibf_gpu_hibf_membership_prepare(hibf_gpu_hibf hibf, hibf_gpu_agent agent, uint64_t max_queries);
int hibf_gpu_hibf_membership_launch(hibf_gpu_hibf hibf,
                                    hibf_gpu_agent agent,
                                    uint64_t const * d_values,
                                    uint64_t const * d_offsets,
                                    uint64_t nq,
                                    uint16_t const * d_thresholds,
                                    uint64_t * d_ids,
                                    uint64_t ids_cap,
                                    uint64_t * d_hit_start,
                                    uint64_t * d_hit_count,
                                    void * stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc): hibf_gpu_hibf_membership_launch without an agent, whose
 * workspace is a 64 MiB array in the library's code object, one per GPU,
 * held until the process ends.  Same inputs and outputs (asynchronous on
 * `stream`, ids in DFS order at d_ids[d_hit_start[q] ..
 * + d_hit_count[q]), ids beyond ids_cap dropped: a query with
 * d_hit_start[q] + d_hit_count[q] > ids_cap lost ids).  At most 65536
 * queries per call (else HIBF_GPU_ERR_INVALID_ARGUMENT); an HIBF whose
 * per-query hit list (4 bytes per run-end technical bin reachable from IBF 0)
 * exceeds the workspace gets HIBF_GPU_ERR_UNSUPPORTED.  Allocates nothing:
 * no hipMalloc, hipFree or allocating library call.  Every call on one GPU
 * shares the workspace, so calls that overlap on a GPU race: serialise them
 * (one stream, or events between streams). */
int hibf_gpu_hibf_membership_launch_noalloc(hibf_gpu_hibf hibf,
                                            uint64_t const * d_values,
                                            uint64_t const * d_offsets,
                                            uint64_t nq,
                                            uint16_t const * d_thresholds,
                                            uint64_t * d_ids,
                                            uint64_t ids_cap,
                                            uint64_t * d_hit_start,
                                            uint64_t * d_hit_count,
                                            void * stream);

/* Serving loop (no counterpart in the reference, whose agents answer one
 * query per call).  Requests q = 0..nq-1 arrive at h_arrival_s[q] seconds
 * (non-decreasing, relative to the call's start) with values
 * h_values[h_offsets[q] .. h_offsets[q+1]) and threshold h_thresholds[q].
 * A request is admitted once it has arrived and its device memory fits
 * budget_bytes; admitted requests are batched until the oldest has waited
 * window_us (or a batch holds max_batch requests / max_batch_values values),
 * copied to the device and answered on one of `streams` streams, so batches
 * overlap.  Outputs: h_latency_s[q] (arrival to ids on the host), up to
 * results_per_query ids per request in h_ids[q*R ..] with the full count in
 * h_counts[q] (ids in the reference's depth-first order; unused id slots are
 * 0).  When h_values is pinned host memory (hipHostMalloc), batches are copied
 * from it directly; otherwise through pinned staging buffers.  Before the
 * first arrival's clock starts, the loop reads a pinned h_values once on the
 * device and runs 64 one-query batches (request 0) per stream, so that no
 * timed batch pays a first-use cost of the copy paths. */
typedef struct hibf_gpu_serve_params
{
    double window_us;
    uint64_t max_batch;
    uint64_t max_batch_values;
    uint64_t budget_bytes;
    uint32_t streams;
    uint32_t results_per_query;
} hibf_gpu_serve_params;

typedef struct hibf_gpu_serve_stats
{
    double seconds;               /* first arrival to last completion */
    uint64_t batches;
    uint64_t lookups;             /* query values checked against one IBF */
    uint64_t row_bytes;
    uint64_t max_in_flight_bytes; /* peak admitted device memory */
    uint64_t overflowed;          /* requests with more than results_per_query ids */
    double longest_pass_s;        /* longest pass of the host loop (collect, admit, launch) */
    double longest_launch_s;      /* longest enqueueing of one batch (copies and launches) */
} hibf_gpu_serve_stats;

int hibf_gpu_hibf_serve(hibf_gpu_hibf hibf,
                        const hibf_gpu_serve_params * params,
                        uint64_t const * h_values,
                        uint64_t const * h_offsets,
                        uint16_t const * h_thresholds,
                        double const * h_arrival_s,
                        uint64_t nq,
                        double * h_latency_s,
                        uint64_t * h_ids,
                        uint32_t * h_counts,
                        hibf_gpu_serve_stats * stats);

/* Online serving (no counterpart in the reference, whose agents answer one
 * query per call): the serving loop above on a host thread of its own, fed by
 * submissions while it runs instead of a trace.  create starts it with the
 * same parameters (batching window, batch limits, admission budget, streams,
 * ids kept per request) and allocates its streams and buffers; submit copies
 * one request (n values and its membership threshold) and returns its ticket,
 * from any number of host threads; the loop admits requests in submission
 * order while their device memory fits the budget, batches them until the
 * oldest has waited the window (or a batch is full) and answers each batch on
 * a free stream.  poll returns finished requests in completion order: up to
 * max_results records and, for record i, its first `kept` ids (the
 * reference's depth-first order) at h_ids[i * results_per_query ..]; it waits
 * up to wait_s seconds when none is ready.  drain waits up to wait_s seconds
 * until every submitted request is answered (HIBF_GPU_ERR_CAPACITY if some
 * are still outstanding then).  destroy answers what was submitted, stops
 * the loop and frees everything; unpolled results are dropped.  Errors:
 * submit returns HIBF_GPU_ERR_BUDGET for a request whose device memory
 * (8 bytes per value, 34 bytes, 8 bytes per kept id) exceeds the budget and
 * HIBF_GPU_ERR_INVALID_ARGUMENT for one longer than max_batch_values; after
 * a device error in the loop, submit, poll and drain return it. */
typedef struct hibf_gpu_server_s * hibf_gpu_server;

typedef struct hibf_gpu_server_result
{
    uint64_t ticket;  /* as returned by hibf_gpu_server_submit                        */
    uint32_t count;   /* ids of the request                                          */
    uint32_t kept;    /* ids in h_ids: min(count, results_per_query), whatever the other
                         requests of its batch found                                 */
    double latency_s; /* submission to result ready                                  */
} hibf_gpu_server_result;

int hibf_gpu_server_create(hibf_gpu_hibf hibf, const hibf_gpu_serve_params * params, hibf_gpu_server * out);
int hibf_gpu_server_submit(hibf_gpu_server server,
                           uint64_t const * h_values,
                           uint64_t n,
                           uint16_t threshold,
                           uint64_t * ticket);
int hibf_gpu_server_poll(hibf_gpu_server server,
                         hibf_gpu_server_result * results,
                         uint64_t * h_ids,
                         uint64_t max_results,
                         double wait_s,
                         uint64_t * n_results);
int hibf_gpu_server_drain(hibf_gpu_server server, double wait_s);
/* Statistics since create (seconds from the first submission to the last answer). */
int hibf_gpu_server_stats(hibf_gpu_server server, hibf_gpu_serve_stats * stats);
int hibf_gpu_server_destroy(hibf_gpu_server server);

/* The serving loop for reads (the tutorial search, test/tutorial/search.cpp:
 * 37-45, served online): request q is the read h_bases[h_read_offsets[q] ..
 * h_read_offsets[q+1]) (characters; dna4 reads anything but C/G/T/U as A);
 * its bases are copied (one byte per base), its k-mer hashes (k in 1..32)
 * and threshold uint16_t(fraction * number of k-mers) are computed on the
 * device (as hibf_gpu_kmer_hash / hibf_gpu_kmer_thresholds), and it is
 * answered like a request of hibf_gpu_hibf_serve.  max_batch_values counts
 * bases; the budget counts bases, hashes and outputs. */
int hibf_gpu_hibf_serve_reads(hibf_gpu_hibf hibf,
                              const hibf_gpu_serve_params * params,
                              const uint8_t * h_bases,
                              const uint64_t * h_read_offsets,
                              uint32_t k,
                              double fraction,
                              const double * h_arrival_s,
                              uint64_t nq,
                              double * h_latency_s,
                              uint64_t * h_ids,
                              uint32_t * h_counts,
                              hibf_gpu_serve_stats * stats);

/* The serving loop for reads hashed to minimisers: as hibf_gpu_hibf_serve_reads,
 * with each read's minimiser hashes (shape, window_size, seed as
 * hibf_gpu_minimiser_hash takes them) computed on the device per batch and the
 * threshold uint16_t(fraction * number of minimisers) taken there; a read's
 * budget counts its bases, hashes (at most one per base) and outputs.
 * Shape and window errors as hibf_gpu_minimiser_hash.  No counterpart in the
 * reference library. */
int hibf_gpu_hibf_serve_reads_minimisers(hibf_gpu_hibf hibf,
                                         const hibf_gpu_serve_params * params,
                                         const uint8_t * h_bases,
                                         const uint64_t * h_read_offsets,
                                         uint64_t shape,
                                         uint32_t window_size,
                                         uint64_t seed,
                                         double fraction,
                                         const double * h_arrival_s,
                                         uint64_t nq,
                                         double * h_latency_s,
                                         uint64_t * h_ids,
                                         uint32_t * h_counts,
                                         hibf_gpu_serve_stats * stats);

/* Batched HIBF counting_agent_type<T>::bulk_count
 * (hierarchical_interleaved_bloom_filter.hpp:462-496, 549-566): d_counts gets
 * nq * number_of_user_bins counters of counter_bytes (1, 2, 4, 8); a user bin's
 * counter is its run sum if that sum >= d_thresholds[q] (size_t, > 0) and 0
 * otherwise; a user bin with several runs in its IBF (cut by deleted bins)
 * keeps the last run sum that reached the threshold, as the reference does.
 * Asynchronous on `stream`.  HIBF_GPU_ERR_UNSUPPORTED for tables that list one
 * user bin in two IBFs (the reference never builds them). */
int hibf_gpu_hibf_bulk_count(hibf_gpu_hibf hibf,
                             hibf_gpu_agent agent,
                             uint64_t const * d_values,
                             uint64_t const * d_offsets,
                             uint64_t nq,
                             uint64_t const * d_thresholds,
                             uint32_t counter_bytes,
                             void * d_counts,
                             void * stream);

/* Host-side table export (for checks and persistence): sizes first, then the
 * arrays hibf_gpu_hibf_create takes.  geoms has n_ibfs entries, word_offsets
 * n_ibfs, tb_offsets n_ibfs+1, next_ibf and user_bin n_tbs. */
int hibf_gpu_hibf_info(hibf_gpu_hibf hibf,
                       uint64_t * n_ibfs,
                       uint64_t * n_tbs,
                       uint64_t * n_words,
                       uint64_t * n_user_bins,
                       uint64_t * depth);
int hibf_gpu_hibf_export_tables(hibf_gpu_hibf hibf,
                                hibf_gpu_ibf_geometry * geoms,
                                uint64_t * word_offsets,
                                uint64_t * tb_offsets,
                                uint64_t * next_ibf,
                                uint64_t * user_bin);

/* ======================================================================== */
/* Reads in: k-mer hashes                                                   */
/* ======================================================================== */

/* k-mer hashes of DNA reads, the values the reference's tutorial search passes
 * to membership_for (test/tutorial/search.cpp:37-41, there computed by
 * seqan3::views::kmer_hash(seqan3::ungapped{k}) over dna4; no counterpart in
 * the library itself).  Read i is d_seq[d_seq_offsets[i] .. d_seq_offsets[i+1])
 * (ASCII).  The hash of a window is the base-4 number of its letters' dna4
 * ranks, first letter most significant (A=0, C=1, G=2, T=U=3, any other
 * character = A); a read shorter than k has none.  Writes n_seqs+1 CSR offsets
 * to d_hash_offsets and the hashes to d_hashes; *h_total is the hash count,
 * HIBF_GPU_ERR_CAPACITY when it exceeds hashes_cap (offsets are written).
 * k in [1, 32].  Synchronous. */
int hibf_gpu_kmer_hash(const uint8_t * d_seq,
                       const uint64_t * d_seq_offsets,
                       uint64_t n_seqs,
                       uint32_t k,
                       uint64_t * d_hash_offsets,
                       uint64_t * d_hashes,
                       uint64_t hashes_cap,
                       uint64_t * h_total,
                       void * stream);

/* Minimiser hashes of DNA reads: seqan3::views::minimiser_hash(shape,
 * window_size, seed) over dna4 (seqan3 3.4.2, the version the reference pins
 * for its tutorial, cmake/package-lock.cmake:50-56; the library itself takes
 * any uint64 values, include/hibf/hierarchical_interleaved_bloom_filter.hpp:
 * 88-93 names minimisers among them).  `shape` is seqan3's binary literal:
 * the most significant set bit is the window's first position, so
 * ungapped{k} is 2^k - 1 and 0b1101 uses positions 0, 1 and 3 of 4; its bit
 * length (the shape size) is at most 32.  Per window position the value is
 * min(forward ^ seed, reverse ^ seed), forward the base-4 number of the dna4
 * ranks at the used positions (first most significant), reverse the same
 * shape over the window's reverse complement.  Of each window of
 * window_size - size + 1 values (all values, when a read has fewer) the
 * minimum is emitted when it changes as seqan3's minimiser view emits it:
 * the first window's rightmost minimum; then the entering value if strictly
 * smaller than the tracked one, or the window's rightmost minimum again when
 * the tracked one leaves.  A read shorter than the shape has none.  Output as
 * hibf_gpu_kmer_hash (CSR offsets, *h_total, HIBF_GPU_ERR_CAPACITY).
 * HIBF_GPU_ERR_INVALID_ARGUMENT for shape 0 or window_size < shape size
 * (seqan3 throws std::invalid_argument there); HIBF_GPU_ERR_UNSUPPORTED for
 * shapes over 32 positions or windows over HIBF_GPU_MAX_MINIMISER_WINDOW
 * values.  No counterpart in the reference library.  Synchronous. */
#define HIBF_GPU_MAX_MINIMISER_WINDOW 4096u
int hibf_gpu_minimiser_hash(const uint8_t * d_seq,
                            const uint64_t * d_seq_offsets,
                            uint64_t n_seqs,
                            uint64_t shape,
                            uint32_t window_size,
                            uint64_t seed,
                            uint64_t * d_hash_offsets,
                            uint64_t * d_hashes,
                            uint64_t hashes_cap,
                            uint64_t * h_total,
                            void * stream);

/* Per-read thresholds uint16_t(fraction * number of k-mers)
 * (test/tutorial/search.cpp:42), truncated and capped at 65535.  Asynchronous. */
int hibf_gpu_kmer_thresholds(const uint64_t * d_hash_offsets,
                             uint64_t n_seqs,
                             double fraction,
                             uint16_t * d_thresholds,
                             void * stream);

/* ======================================================================== */
/* Build side: HyperLogLog sketches                                         */
/* ======================================================================== */

/* seqan::hibf::sketch::hyperloglog(bits) + add() over n value sets on the
 * device: sketch::compute_sketches (src/sketch/compute_sketches.cpp:27-43),
 * one sketch per user bin, with hyperloglog::add (src/sketch/hyperloglog.cpp:
 * 48-64).  Set i is d_values[d_offsets[i] .. d_offsets[i + 1]); its 2^bits
 * uint8 registers are written to d_registers + (i << bits).  bits in [5, 32]
 * (the constructor's range and message, hyperloglog.cpp:28-29, else
 * HIBF_GPU_ERR_INVALID_ARGUMENT).  Synchronous. */
int hibf_gpu_hll_sketch(const uint64_t * d_values,
                        const uint64_t * d_offsets,
                        uint64_t n_sketches,
                        uint32_t bits,
                        uint8_t * d_registers,
                        void * stream);

/* As hibf_gpu_hll_sketch for the synthetic user bins of
 * hibf_gpu_hibf_build_synthetic: set u is the h_sizes[u] values of the
 * library's seeded generator (no counterpart in the reference).  Synchronous. */
int hibf_gpu_hll_sketch_synthetic(uint64_t seed,
                                  const uint64_t * h_sizes,
                                  uint64_t n_sketches,
                                  uint32_t bits,
                                  uint8_t * d_registers,
                                  void * stream);

/* hyperloglog::merge (hyperloglog.cpp:136-149) over groups: group g's
 * registers (d_out + (g << bits)) are the register-wise max of the sketches
 * d_members[d_group_offsets[g] .. d_group_offsets[g + 1]).  Asynchronous. */
int hibf_gpu_hll_merge(const uint8_t * d_registers,
                       uint32_t bits,
                       const uint64_t * d_members,
                       const uint64_t * d_group_offsets,
                       uint64_t n_groups,
                       uint8_t * d_out,
                       void * stream);

/* hyperloglog::estimate (hyperloglog.cpp:66-129) of n sketches into
 * h_estimates (host).  The float sum runs in the reference's eight lanes and
 * lane order on the device; the division and the small-value correction run
 * on the host in double.  Synchronous. */
int hibf_gpu_hll_estimate(const uint8_t * d_registers, uint64_t n_sketches, uint32_t bits, double * h_estimates,
                          void * stream);

/* The layout DP's union estimates (sketch::toolbox::precompute_union_estimates_for,
 * src/sketch/toolbox.cpp:35-51, for every column at once): for all
 * 0 <= k < j < n_positions, the estimate of the union of the sketches
 * d_positions[k] .. d_positions[j] (merge_and_estimate, truncated to uint64),
 * bit-identical to the reference's.  Column j's j values start at
 * d_out[j (j - 1) / 2]; d_out holds n_positions (n_positions - 1) / 2 values.
 * Sketches of 5 to 12 bits (else HIBF_GPU_ERR_UNSUPPORTED).  Synchronous. */
int hibf_gpu_hll_union_estimates(const uint8_t * d_registers,
                                 uint64_t n_sketches,
                                 uint32_t bits,
                                 const uint64_t * d_positions,
                                 uint64_t n_positions,
                                 uint64_t * d_out,
                                 void * stream);

/* ======================================================================== */
/* Archives: cereal BinaryArchive layout                                    */
/* ======================================================================== */

/* interleaved_bloom_filter::serialize (interleaved_bloom_filter.hpp:471-485)
 * with bit_vector::save/load (bit_vector.hpp:885-924), in the byte layout a
 * cereal::BinaryOutputArchive writes for it: uint32 version (1); uint64 bins,
 * technical_bins, bin_size, hash_shift, bin_words, hash_funs; uint64 bit count
 * followed by the bit_vector words; uint64 occupancy length followed by the
 * occupancy counters; one byte track_occupancy.  Files written by util/inspect-
 * style tools (a bare archive of the object) load here and vice versa.
 * save: append != 0 appends to an existing file (for wrapper formats).
 * load: reads from byte `offset` of the file and reports the offset after the
 * object in *end_offset (may be NULL).  Errors: HIBF_GPU_ERR_IO. */
int hibf_gpu_ibf_save(hibf_gpu_ibf ibf, const char * path, int append);
int hibf_gpu_ibf_load(const char * path, uint64_t offset, hibf_gpu_ibf * out, uint64_t * end_offset);

/* hierarchical_interleaved_bloom_filter::serialize
 * (hierarchical_interleaved_bloom_filter.hpp:267-277): uint32 version (1);
 * uint64 number_of_user_bins; ibf_vector (uint64 count, then each IBF as
 * above); next_ibf_id (uint64 count, then per IBF a uint64 length and the
 * ids); prev_ibf_id (uint64 count, then per IBF uint64 ibf_idx, bin_idx);
 * ibf_bin_to_user_bin_id (as next_ibf_id).  The device HIBF keeps neither
 * occupancy nor prev_ibf_id: save writes zero occupancy with track_occupancy
 * false (the reference's HIBF build at empty_bin_fraction 0) and prev_ibf_id
 * derived from the merged bins ({0, 0} for the root); load checks and drops
 * them. */
int hibf_gpu_hibf_save(hibf_gpu_hibf hibf, const char * path, int append);
int hibf_gpu_hibf_load(const char * path, uint64_t offset, hibf_gpu_hibf * out, uint64_t * end_offset);

/* ======================================================================== */
/* HIBF construction (host layout + device bulk emplace)                    */
/* ======================================================================== */

/* Numeric fields of seqan::hibf::config (config.hpp:111-324); the reference's
 * input_fn is replaced by the user-bin contents passed to the build calls. */
typedef struct hibf_gpu_config
{
    uint64_t number_of_user_bins;
    uint64_t number_of_hash_functions; /* default 2    */
    double maximum_fpr;                /* default 0.05 */
    double relaxed_fpr;                /* default 0.3  */
    uint64_t threads;                  /* default 1    */
    uint64_t sketch_bits;              /* default 12   */
    uint64_t tmax;                     /* 0 = next_multiple_of_64(ceil(sqrt(#UB))) */
    double empty_bin_fraction;         /* default 0    */
    double alpha;                      /* default 1.2  */
    double max_rearrangement_ratio;    /* default 0.5  */
    int32_t disable_estimate_union;
    int32_t disable_rearrangement;
    int32_t track_occupancy;
    int32_t reserved;
} hibf_gpu_config;

/* The reference's default member initialisers (config.hpp:136-324). */
int hibf_gpu_config_default(hibf_gpu_config * c);
/* Replaces config::validate_and_set_defaults (src/config.cpp:64-130):
 * HIBF_GPU_ERR_CONFIG with the reference's message (hibf_gpu_last_error). */
int hibf_gpu_config_validate(hibf_gpu_config * c);

/* build::bin_size_in_bits (include/hibf/build/bin_size_in_bits.hpp:35-45):
 * bits per technical bin for `elements` values at false-positive rate fpr. */
uint64_t hibf_gpu_bin_size_in_bits(double fpr, uint64_t hash_count, uint64_t elements);

/* Replaces layout::compute_layout (src/layout/compute_layout.cpp:26-70) for
 * the given user-bin cardinalities, with merged-bin cardinalities taken as
 * sums and no similarity rearrangement.  Writes the reference's layout text
 * format (src/layout/layout.cpp:125-136); HIBF_GPU_ERR_CAPACITY + *needed
 * when cap is too small. */
int hibf_gpu_layout_compute(const hibf_gpu_config * c,
                            uint64_t const * h_kmer_counts,
                            char * text,
                            uint64_t cap,
                            uint64_t * needed);

/* compute_layout(config, kmer_counts, sketches) (layout/compute_layout.cpp:26-70):
 * as hibf_gpu_layout_compute, but with the user bins' HyperLogLog sketches
 * (d_registers: number_of_user_bins x 2^sketch_bits uint8, device) the DP
 * uses union estimates of merged bins and, unless disable_rearrangement,
 * first rearranges each level's user bins by similarity
 * (sketch::toolbox::rearrange_bins), as the reference.  The union and pair
 * estimates run on the device for sketches of at most 12 bits.  With
 * disable_estimate_union the sketches are not used.  A user bin with an empty
 * sketch makes the rearrangement fail (HIBF_GPU_ERR_INVALID_ARGUMENT), as in
 * the reference.  h_times (may be null) gets the seconds spent in union
 * estimation [0] and rearrangement [1]. */
int hibf_gpu_layout_compute_sketched(const hibf_gpu_config * config,
                                     const uint64_t * h_kmer_counts,
                                     const uint8_t * d_registers,
                                     char * text,
                                     uint64_t cap,
                                     uint64_t * needed,
                                     double * h_times);

/* compute_layout(config, kmer_counts, sketches, positions, timers)
 * (layout/compute_layout.cpp:26-55) with the sketches in host memory
 * (h_registers: number_of_user_bins x 2^sketch_bits uint8; uploaded for the
 * call) and, when h_positions is not null, only the n_positions user bins it
 * lists laid out (each < number_of_user_bins).  Otherwise as
 * hibf_gpu_layout_compute_sketched.  Synchronous. */
int hibf_gpu_layout_compute_host_sketches(const hibf_gpu_config * config,
                                          const uint64_t * h_kmer_counts,
                                          const uint8_t * h_registers,
                                          const uint64_t * h_positions,
                                          uint64_t n_positions,
                                          char * text,
                                          uint64_t cap,
                                          uint64_t * needed,
                                          double * h_times);

/* Replaces hierarchical_interleaved_bloom_filter(config&) and (config&, layout)
 * (src/hierarchical_interleaved_bloom_filter.cpp:211-253): user bin u holds
 * h_values[h_ub_offsets[u] .. h_ub_offsets[u+1]) in the order the reference's
 * input_fn would insert them (duplicates ignored).  The layout is computed
 * unless layout_text (reference layout format) is given.  IBF geometries and
 * tables follow the reference's hierarchical build; the bit matrices are
 * filled by the device bulk emplace.  A user bin split over several technical
 * bins is chunked in the iteration order the reference's k-mer set
 * (robin_hood::unordered_flat_set<uint64_t>) has after the same insertions,
 * clears and merges (src/build/insert_into_ibf.cpp:28-47), replayed on the
 * host; the tuning key "build.split_order" = 0 chunks by sorted value ranges
 * instead (faster, matrices differ from the reference's). */
int hibf_gpu_hibf_build(const hibf_gpu_config * c,
                        uint64_t const * h_ub_offsets,
                        uint64_t const * h_values,
                        const char * layout_text,
                        hibf_gpu_hibf * out);

/* As hibf_gpu_hibf_build for synthetic user bins: user bin u holds
 * hibf_gpu_synth_value(seed, u, i) for i < h_ub_sizes[u] (pairwise distinct),
 * inserted in order of i. */
int hibf_gpu_hibf_build_synthetic(const hibf_gpu_config * c,
                                  uint64_t const * h_ub_sizes,
                                  uint64_t seed,
                                  const char * layout_text,
                                  hibf_gpu_hibf * out);

/* Dry run of a build (host only, no device memory): number of IBFs, arena
 * words, emplace segments and (value, technical bin) insertions the build of
 * these user-bin sizes would make.  For capacity planning. */
int hibf_gpu_hibf_plan_stats(const hibf_gpu_config * c,
                             uint64_t const * h_ub_sizes,
                             const char * layout_text,
                             uint64_t * n_ibfs,
                             uint64_t * n_words,
                             uint64_t * n_segments,
                             uint64_t * n_emplaces);

#ifdef __cplusplus
}
#endif

#endif /* HIBF_GPU_H */
