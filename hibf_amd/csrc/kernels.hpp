// Host-side launch wrappers of the HIBF query kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"

namespace hibf_amd
{

// LDS the descent kernels may give a block's per-IBF counters and per-level
// event lists ((1 + depth) x largest IBF x 4 bytes).  HIBFs that need more
// (large IBFs, deep trees) keep that state in a per-block region of device
// memory instead (HibfLaunch::gstate_words).
constexpr uint64_t kHibfLdsStateBytes = 64u * 1024u;
// Deepest HIBF (levels) the descent kernel handles.
constexpr uint32_t kHibfMaxDepth = 16;

// Kinds of a technical bin in the descent tables (hierarchical_interleaved_bloom_filter.hpp:329-354).
enum TbKind : uint32_t
{
    kTbRegular = 0u, // user bin, not the last bin of its run
    kTbRunEnd = 1u,  // user bin, last bin of its run (or last bin of the IBF)
    kTbMerged = 2u,  // merged bin: payload = child IBF index
    kTbDeleted = 3u  // deleted bin
};

// One IBF of an HIBF as the descent kernel sees it.
struct HibfIbfDesc
{
    uint64_t word_offset; // first word in the arena
    uint64_t bin_size;
    uint64_t tb_offset; // first entry in tb_info
    uint32_t bins;
    uint32_t bin_words;
    uint32_t hash_shift;
    uint32_t hash_funs;
    uint32_t lpv_log; // count_into_lds_dispatch shape
    uint32_t pad;
};

void count_shape(uint32_t bin_words, uint32_t & lpv_log, uint32_t & cw, uint32_t & tile_words);

// An IBF's descriptor at a uniform address, with scalar loads (the tables do
// not change while a kernel runs).
__device__ __forceinline__ HibfIbfDesc load_desc(HibfIbfDesc const * p)
{
    static_assert(sizeof(HibfIbfDesc) == 48, "six 64-bit words");
    uint64_t const * const w = reinterpret_cast<uint64_t const *>(p);
    uint64_t raw[6];
#pragma unroll
    for (int i = 0; i < 6; ++i)
        raw[i] = const_load(w + i);
    HibfIbfDesc d;
    __builtin_memcpy(&d, raw, sizeof d);
    return d;
}

// Launch-shape knobs of the count kernels, chosen by the autotuner
// (hibf_gpu_tune_set); defaults are the tuned values for MI355X.
struct CountTuning
{
    uint32_t wide_kernel = 1; // use count_wide_kernel for even bin_words >= 64
    uint32_t wide_unroll = 2; // values in flight per wave (1, 2, 4, 8)
    uint32_t contains_unroll = 4; // bulk_contains: values whose rows a wave loads at once (1, 2, 4, 8)
    uint32_t wide_waves = 4;  // waves per block (1..4)
    uint32_t hibf_wave = 0;   // HIBF descent: one wave per query (1) or one block per query (0)
    uint32_t nt_loads = 0;    // HIBF descent: non-temporal row loads
    uint32_t counter_bits = 16; // count_wide_kernel LDS counter width (8 or 16)
    uint32_t row_per_lane = 0;  // HIBF descent: one value and its whole row per lane when every row is <= 4 words
    uint32_t sliced_counts = 1; // HIBF descent over rows <= 4 words: bit-sliced per-lane counts
    uint32_t nt_stores = 1;     // count_wide_kernel and the HIBF counting zero fill: non-temporal (streaming) stores
    uint32_t wide_nt_loads = 1; // count_wide_kernel with 2 or 3 hash functions: non-temporal row loads
    uint32_t hibf_block_waves = 4; // HIBF membership, query per block: waves per block (1, 2, 4)
    uint32_t hibf_unroll = 4;      // HIBF membership over rows <= 4 words, sliced counts: value groups in flight (2, 4)
    uint32_t hibf_stream_kib = 1u << 16; // HIBF descent over rows <= 4 words: IBFs of at least this many KiB
                                         // (default 64 MiB) take non-temporal row loads (0: none)
    uint32_t split_order = 1;      // HIBF build: split user bins chunked in the reference's set order (1) or by value ranges (0)
};
extern CountTuning g_count_tuning;

// Largest single device allocation the library makes, in bytes (0: no
// limit).  Set through hibf_gpu_tune_set("alloc.max_bytes"); a larger request
// fails as hipMalloc fails when memory runs out, so the out-of-memory paths of
// every entry point can be exercised.
extern uint64_t g_alloc_max_bytes;

// Every device allocation of the library goes through here.
inline hipError_t device_malloc(void ** p, size_t bytes)
{
    if (g_alloc_max_bytes != 0 && bytes > g_alloc_max_bytes)
    {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipMalloc(p, bytes);
}
template <typename T>
hipError_t device_malloc(T ** p, size_t bytes)
{
    void * v = nullptr;
    hipError_t const e = device_malloc(&v, bytes);
    *p = static_cast<T *>(v);
    return e;
}

struct IbfView;
bool count_wide_applies(IbfView const & g);
hipError_t launch_count_wide(IbfView const & g,
                             uint64_t const * values,
                             uint64_t const * offsets,
                             uint64_t nq,
                             uint32_t counter_bytes,
                             void * counts,
                             unsigned max_blocks,
                             hipStream_t stream);

hipError_t launch_emplace(uint64_t * words,
                          uint64_t bin_size,
                          uint32_t bin_words,
                          uint32_t hash_shift,
                          uint32_t hash_funs,
                          uint64_t const * values,
                          uint64_t const * bins,
                          uint64_t bin,
                          uint64_t n,
                          uint64_t bin_count,
                          hipStream_t stream);

hipError_t launch_emplace_synth(uint64_t * words,
                                uint64_t bin_size,
                                uint32_t bin_words,
                                uint32_t hash_shift,
                                uint32_t hash_funs,
                                uint64_t seed,
                                uint64_t first_bin,
                                uint64_t n_bins,
                                uint64_t per_bin,
                                hipStream_t stream);

hipError_t launch_restride(uint64_t const * old_words,
                           uint64_t old_bw,
                           uint64_t * new_words,
                           uint64_t new_bw,
                           uint64_t bin_size,
                           hipStream_t stream);
hipError_t launch_clear(uint64_t * words,
                        uint64_t bin_size,
                        uint32_t bin_words,
                        uint64_t const * d_bins,
                        uint32_t nbins,
                        hipStream_t stream);

hipError_t launch_contains(IbfView const & g, uint64_t const * values, uint64_t n, uint64_t * out, hipStream_t stream);

hipError_t launch_count(IbfView const & g,
                        uint64_t const * values,
                        uint64_t const * offsets,
                        uint64_t nq,
                        uint32_t counter_bytes,
                        void * counts,
                        unsigned max_blocks,
                        hipStream_t stream);

// IBF membership_for from counts already in device memory: mask bit b of query
// q (mask words q * mask_words ..) = counts[q * bins + b] >= thresholds[q], and
// hit_count[q] = the number of set bits; then the set bins of each mask, in
// ascending order, into ids[hit_offsets[q] ..) (positions >= ids_cap dropped).
hipError_t launch_threshold_mask(uint16_t const * counts,
                                 uint64_t bins,
                                 uint64_t nq,
                                 uint16_t const * thresholds,
                                 uint64_t * hit_count,
                                 uint64_t * mask,
                                 uint64_t mask_words,
                                 hipStream_t stream);
hipError_t launch_mask_compact(uint64_t const * mask,
                               uint64_t mask_words,
                               uint64_t nq,
                               uint64_t const * hit_offsets,
                               uint64_t * ids,
                               uint64_t ids_cap,
                               hipStream_t stream);

struct HibfLaunch
{
    HibfIbfDesc const * ibfs;
    uint64_t const * arena;
    uint64_t const * tb_info; // (kind << 32) | payload
    uint32_t max_bins;        // largest IBF of this HIBF
    uint32_t depth;           // levels
    uint64_t n_user_bins;
    uint64_t max_hits;        // most ids one query can produce (run-end TBs reachable from IBF 0)
    uint32_t max_bin_words;   // widest row of this HIBF (words)
    uint32_t uniform_hash_funs; // hash functions of every IBF, 0 if they differ
    uint32_t nt_stores;       // counting: non-temporal stores of the zero fill
    // Membership output form.  0: query q's ids go to ids[hit_start[q] ..) with
    // hit_start taken in atomic order, so one query may use many slots.  C > 0:
    // query q writes its first min(n, C) ids to ids[q * C ..) (hit_start[q] =
    // q * C) and hit_count[q] = n, whatever the other queries found.
    uint32_t per_query_cap;
    // 0: a block's counters and event lists are in LDS.  Otherwise they are
    // (1 + depth) x max_bins words of device memory per block, after the
    // block's hit list in the scratch the launch is given.
    uint64_t gstate_words;
    // IBFs of at least this many bytes are read with non-temporal (streaming)
    // row loads by the narrow-row kernels: too large for the caches to keep
    // their rows between queries.  ~0: never.
    uint64_t stream_bytes;
    // Agent counters (device): [0] ids found, [1] next query (work stealing),
    // [2] IBF visits, [3] row bytes fetched (h * bin_words * 8 per value and
    // visit), [4] lookups (query values checked against one IBF), [5] set when
    // ids were dropped because the output was full, [6] length of the overflow
    // query list, [7] next entry of that list.
    unsigned long long * counters;
};

// Zeroes the agent counters selected by `mask` (bit i = counter i).  A kernel,
// so it is safe inside captured graphs.
hipError_t launch_reset_counters(unsigned long long * counters, uint32_t mask, hipStream_t stream);

// Membership descent, one block per query.  Writes, per query,
// hit_start/hit_count into ids_tmp (atomic order of queries).  With a query
// list it answers only the queries listed there (counters[6] entries).
// scratch holds, per block, max_hits words of hit list and then, when
// h.gstate_words != 0, that many words of counters and event lists.
hipError_t launch_hibf_membership(HibfLaunch const & h,
                                  uint64_t const * values,
                                  uint64_t const * offsets,
                                  uint64_t nq,
                                  uint16_t const * thresholds,
                                  uint32_t * scratch,
                                  unsigned blocks,
                                  uint64_t * ids_tmp,
                                  uint64_t ids_cap,
                                  uint64_t * hit_start,
                                  uint64_t * hit_count,
                                  uint64_t const * qlist,
                                  hipStream_t stream);

// One wave per query (hibf_wave.hip); queries with more ids than the wave's
// LDS buffer are appended to `overflow` for launch_hibf_membership.
size_t hibf_wave_lds_bytes(HibfLaunch const & h, unsigned waves_per_block);
hipError_t launch_hibf_wave_membership(HibfLaunch const & h,
                                       uint64_t const * values,
                                       uint64_t const * offsets,
                                       uint64_t nq,
                                       uint16_t const * thresholds,
                                       unsigned blocks,
                                       uint64_t * ids,
                                       uint64_t ids_cap,
                                       uint64_t * hit_start,
                                       uint64_t * hit_count,
                                       uint64_t * overflow,
                                       hipStream_t stream);
hipError_t launch_hibf_wave_count(HibfLaunch const & h,
                                  uint64_t const * values,
                                  uint64_t const * offsets,
                                  uint64_t nq,
                                  uint64_t const * thresholds,
                                  uint32_t counter_bytes,
                                  void * counts,
                                  unsigned blocks,
                                  hipStream_t stream);

// Counting descent, one block per query; scratch holds gstate_words words per
// block when h.gstate_words != 0 (unused otherwise).
hipError_t launch_hibf_count(HibfLaunch const & h,
                             uint64_t const * values,
                             uint64_t const * offsets,
                             uint64_t nq,
                             uint64_t const * thresholds,
                             uint32_t counter_bytes,
                             void * counts,
                             uint32_t * scratch,
                             unsigned blocks,
                             hipStream_t stream);

// Device form of a build-plan segment (same layout as host::Segment): values
// [begin, begin+count) of user bin `ub` go to technical bin `tb` of IBF `ibf`;
// with bit 0 of flags set (a split user bin chunked in the reference's set
// order), value i of the user bin goes to technical bin tb + chunk_ids[begin + i].
struct EmplaceSegment
{
    uint32_t ibf;
    uint32_t tb;
    uint32_t ub;
    uint32_t flags;
    uint64_t begin;
    uint64_t count;
};

// A block's share of one segment: values [start, start+len) of segment seg.
struct EmplaceChunk
{
    uint32_t seg;
    uint32_t len;
    uint64_t start;
};

// Executes a build plan: sets the bits of every (value, technical bin) pair of
// every segment.  Values come from values[ub_offsets[ub] + i] or, when
// values == nullptr, from synth_value(seed, ub, i).  chunk_ids (entries of
// chunk_id_bytes, 1 or 2) serve the chunked segments.
hipError_t launch_segment_emplace(HibfIbfDesc const * desc,
                                  uint64_t * arena,
                                  EmplaceSegment const * segs,
                                  EmplaceChunk const * chunks,
                                  uint64_t n_chunks,
                                  uint64_t const * values,
                                  uint64_t const * ub_offsets,
                                  uint64_t seed,
                                  uint8_t const * chunk_ids,
                                  uint32_t chunk_id_bytes,
                                  hipStream_t stream);

// out[offsets[q] + i] = src[start[q] + i] for i < count[q].
hipError_t launch_gather_hits(uint64_t const * src,
                              uint64_t const * start,
                              uint64_t const * count,
                              uint64_t const * offsets,
                              uint64_t nq,
                              uint64_t * out,
                              hipStream_t stream);

// With b = base + (*d_base, or 0 when d_base is null): out[0] = b (only when
// write_first), out[1 + i] = b + incl[i].  d_base may point at out[0].
hipError_t launch_finalize_offsets(uint64_t const * incl, uint64_t n, uint64_t base, uint64_t const * d_base,
                                   bool write_first, uint64_t * out, hipStream_t stream);

} // namespace hibf_amd

namespace hibf_amd
{
// hipCUB inclusive sum; call with temp == nullptr to size temp_bytes.
hipError_t inclusive_sum_u64(void * temp, size_t & temp_bytes, uint64_t const * in, uint64_t * out, uint64_t n,
                             hipStream_t stream);
// hipCUB stable radix sort of (key, value) pairs on key bits [0, end_bit).
hipError_t radix_sort_pairs_u64(void * temp, size_t & temp_bytes, uint64_t const * keys_in, uint64_t * keys_out,
                                uint64_t const * vals_in, uint64_t * vals_out, uint64_t n, int end_bit,
                                hipStream_t stream);
// Steps of the emplace with track_occupancy (ibf_kernels.hip): the bit
// positions of the pairs (key = row * technical_bins + bin, `invalid_key`
// for bins >= bin_count; value = pair index), and, after a stable sort by
// key, a mark on each pair that is the first to touch a position unset in
// words, then occupancy[bin] += 1 per marked pair.
hipError_t launch_emplace_positions(uint64_t bin_size, uint32_t bin_words, uint32_t hash_shift, uint32_t hash_funs,
                                    uint64_t const * values, uint64_t const * bins, uint64_t bin, uint64_t n,
                                    uint64_t bin_count, uint64_t invalid_key, uint64_t * keys, uint64_t * owners,
                                    hipStream_t stream);
hipError_t launch_mark_first_touches(uint64_t const * words, uint64_t const * keys, uint64_t const * owners,
                                     uint64_t m, uint64_t invalid_key, uint8_t * marked, hipStream_t stream);
hipError_t launch_count_marked(uint8_t const * marked, uint64_t const * bins, uint64_t bin, uint64_t n,
                               unsigned long long * occupancy, hipStream_t stream);
// HyperLogLog sketches (hll.hip): registers of the synthetic user bins
// (synth_value(seed, u, i), i < h_sizes[u]) and the estimates of n sketches on
// the device (synchronous).
int hll_sketch_synthetic(uint64_t seed, uint64_t const * h_sizes, uint64_t n, uint32_t bits, uint8_t * d_registers,
                         hipStream_t s);
int hll_estimates(uint8_t const * d_registers, uint64_t n, uint32_t bits, double * h_estimates, hipStream_t s);

// k-mer hashes of reads on the device without host waits (kmer.hip).
// Minimiser hashes (kmer.hip): shape/window validation as hibf_gpu_minimiser_hash
// (a status), the offsets scan's temporary bytes for n reads, and the
// asynchronous count / scan / thresholds / write pipeline.
int minimiser_check(uint64_t shape, uint32_t window_size);
hipError_t minimiser_scan_bytes(uint64_t n, size_t & bytes);
hipError_t minimiser_hash_async(uint8_t const * d_seq,
                                uint64_t const * d_seq_offsets,
                                uint64_t n,
                                uint64_t shape,
                                uint32_t window_size,
                                uint64_t seed,
                                uint64_t * d_hash_offsets,
                                uint64_t * d_hashes,
                                void * scan_tmp,
                                size_t scan_bytes,
                                double fraction,
                                uint16_t * d_thresholds,
                                hipStream_t s);
hipError_t kmer_hash_given_offsets_async(uint8_t const * d_seq,
                                         uint64_t const * d_seq_offsets,
                                         uint64_t n,
                                         uint32_t k,
                                         uint64_t const * d_hash_offsets,
                                         uint64_t * d_hashes,
                                         hipStream_t s);

} // namespace hibf_amd
