// IBF kernels for gfx950: bulk emplace, clear, bulk_contains and bulk_count.
#include "device_common.hpp"
#include "kernels.hpp"

namespace hibf_amd
{

// One thread per (value, bin) pair: set bit hash_and_fit(v, seed_k) + bin for
// every k < h (src/interleaved_bloom_filter.cpp:119-140).  atomicOr makes the
// result independent of order, so the matrix equals the sequential one.
__global__ void __launch_bounds__(256) emplace_kernel(uint64_t * __restrict__ words,
                                                      uint64_t bin_size,
                                                      uint32_t bin_words,
                                                      uint32_t hash_shift,
                                                      uint32_t hash_funs,
                                                      uint64_t const * __restrict__ values,
                                                      uint64_t const * __restrict__ bins,
                                                      uint64_t bin,
                                                      uint64_t n,
                                                      uint64_t bin_count)
{
    uint64_t const stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    {
        uint64_t const v = values[i];
        uint64_t const b = bins ? bins[i] : bin;
        if (b >= bin_count) // the reference asserts bin < bins; out-of-range pairs are skipped
            continue;
        uint64_t const word_in_row = b >> 6;
        unsigned long long const bit = 1ULL << (b & 63u);
#pragma unroll
        for (int k = 0; k < 5; ++k)
            if (static_cast<uint32_t>(k) < hash_funs)
            {
                uint64_t const w = row_word(v, hash_seed(k), hash_shift, bin_size, bin_words) + word_in_row;
                atomicOr(reinterpret_cast<unsigned long long *>(words + w), bit);
            }
    }
}

// Emplace with track_occupancy (src/interleaved_bloom_filter.cpp:119-140),
// pairs i = 0, 1, ... in order: the reference reads each hash's bit before
// setting it and counts the value iff some read found an unset bit.  Every
// bit belongs to one bin, so that is the case iff the value is the first pair
// of the call to touch one of its positions and that position was unset
// before the call.  Step 1: key[i * h + k] = position of hash k of pair i
// (row * technical_bins + bin, the bit index in the matrix), owner = i.
__global__ void __launch_bounds__(256) emplace_positions_kernel(uint64_t bin_size,
                                                                uint32_t bin_words,
                                                                uint32_t hash_shift,
                                                                uint32_t hash_funs,
                                                                uint64_t const * __restrict__ values,
                                                                uint64_t const * __restrict__ bins,
                                                                uint64_t bin,
                                                                uint64_t n,
                                                                uint64_t bin_count,
                                                                uint64_t invalid_key,
                                                                uint64_t * __restrict__ keys,
                                                                uint64_t * __restrict__ owners)
{
    uint64_t const stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    {
        uint64_t const v = values[i];
        uint64_t const b = bins ? bins[i] : bin;
        for (uint32_t k = 0; k < hash_funs; ++k)
        {
            keys[i * hash_funs + k] = b < bin_count
                                          ? row_word(v, hash_seed(static_cast<int>(k)), hash_shift, bin_size, bin_words) *
                                                    64u +
                                                b
                                          : invalid_key;
            owners[i * hash_funs + k] = i;
        }
    }
}

// Step 2, over the pairs sorted stably by key (equal keys keep pair order, so
// a key's first entry is its first touch): mark the owner of a first touch of
// a position whose bit is unset in words (read before the emplace).
__global__ void __launch_bounds__(256) mark_first_touches_kernel(uint64_t const * __restrict__ words,
                                                                 uint64_t const * __restrict__ keys,
                                                                 uint64_t const * __restrict__ owners,
                                                                 uint64_t m,
                                                                 uint64_t invalid_key,
                                                                 uint8_t * __restrict__ marked)
{
    uint64_t const stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < m; e += stride)
    {
        uint64_t const key = keys[e];
        if (key == invalid_key || (e > 0 && keys[e - 1] == key))
            continue;
        if (((words[key >> 6] >> (key & 63u)) & 1ULL) == 0ULL)
            marked[owners[e]] = 1u;
    }
}

// Step 3 (after the emplace): occupancy[bin of pair i] += 1 per marked pair.
__global__ void __launch_bounds__(256) count_marked_kernel(uint8_t const * __restrict__ marked,
                                                           uint64_t const * __restrict__ bins,
                                                           uint64_t bin,
                                                           uint64_t n,
                                                           unsigned long long * __restrict__ occupancy)
{
    uint64_t const stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        if (marked[i])
            atomicAdd(&occupancy[bins ? bins[i] : bin], 1ULL);
}

// Generator-driven bulk emplace: values synth_value(seed, first_bin + b, j)
// for j < per_bin go to technical bin first_bin + b, b < n_bins.
__global__ void __launch_bounds__(256) emplace_synth_kernel(uint64_t * __restrict__ words,
                                                            uint64_t bin_size,
                                                            uint32_t bin_words,
                                                            uint32_t hash_shift,
                                                            uint32_t hash_funs,
                                                            uint64_t seed,
                                                            uint64_t first_bin,
                                                            uint64_t per_bin,
                                                            uint64_t n)
{
    uint64_t const stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    {
        uint64_t const b = first_bin + i / per_bin;
        uint64_t const v = synth_value(seed, b, i % per_bin);
        uint64_t const word_in_row = b >> 6;
        unsigned long long const bit = 1ULL << (b & 63u);
#pragma unroll
        for (int k = 0; k < 5; ++k)
            if (static_cast<uint32_t>(k) < hash_funs)
            {
                uint64_t const w = row_word(v, hash_seed(k), hash_shift, bin_size, bin_words) + word_in_row;
                atomicOr(reinterpret_cast<unsigned long long *>(words + w), bit);
            }
    }
}

// Build-plan execution: one block per chunk of one segment.
__global__ void __launch_bounds__(256) segment_emplace_kernel(HibfIbfDesc const * __restrict__ desc,
                                                              uint64_t * __restrict__ arena,
                                                              EmplaceSegment const * __restrict__ segs,
                                                              EmplaceChunk const * __restrict__ chunks,
                                                              uint64_t n_chunks,
                                                              uint64_t const * __restrict__ values,
                                                              uint64_t const * __restrict__ ub_offsets,
                                                              uint64_t seed,
                                                              uint8_t const * __restrict__ chunk_ids,
                                                              uint32_t chunk_id_bytes)
{
    for (uint64_t c = blockIdx.x; c < n_chunks; c += gridDim.x)
    {
        EmplaceChunk const ch = chunks[c];
        EmplaceSegment const sg = segs[ch.seg];
        HibfIbfDesc const d = desc[sg.ibf];
        uint64_t * const words = arena + d.word_offset;
        bool const chunked = (sg.flags & 1u) != 0u;
        uint64_t const base = values ? ub_offsets[sg.ub] : 0ULL;
        for (uint32_t t = threadIdx.x; t < ch.len; t += blockDim.x)
        {
            // chunked: value i of the user bin, its technical bin from its chunk number
            uint64_t const i = chunked ? ch.start + t : sg.begin + ch.start + t;
            uint32_t tb = sg.tb;
            if (chunked)
                tb += chunk_id_bytes == 1u ? chunk_ids[sg.begin + i]
                                           : reinterpret_cast<uint16_t const *>(chunk_ids)[sg.begin + i];
            uint64_t const word_in_row = tb >> 6;
            unsigned long long const bit = 1ULL << (tb & 63u);
            uint64_t const v = values ? values[base + i] : synth_value(seed, sg.ub, i);
#pragma unroll
            for (int k = 0; k < 5; ++k)
                if (static_cast<uint32_t>(k) < d.hash_funs)
                {
                    uint64_t const w = row_word(v, hash_seed(k), d.hash_shift, d.bin_size, d.bin_words) + word_in_row;
                    atomicOr(reinterpret_cast<unsigned long long *>(words + w), bit);
                }
        }
    }
}

// clear(bin): zero the bin's column in every row (src/interleaved_bloom_filter.cpp:142-148).
__global__ void __launch_bounds__(256) clear_kernel(uint64_t * __restrict__ words,
                                                    uint64_t bin_size,
                                                    uint32_t bin_words,
                                                    uint64_t const * __restrict__ bins,
                                                    uint32_t nbins)
{
    uint64_t const stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t r = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < bin_size; r += stride)
    {
        for (uint32_t j = 0; j < nbins; ++j)
        {
            uint64_t const b = bins[j];
            unsigned long long const keep = ~(1ULL << (b & 63u));
            atomicAnd(reinterpret_cast<unsigned long long *>(words + r * bin_words + (b >> 6)), keep);
        }
    }
}

// Batched bulk_contains (src/interleaved_bloom_filter.cpp:204-291): one wave per
// 64 values.  Lane j hashes value j; the wave then walks the values U at a
// time, broadcasting their row starts (readlane), issuing the h row loads of
// all U values before it ANDs and stores them, so U x h row fetches are in
// flight per wave.  With vec set (even row width, 16-byte aligned words and
// output) a lane moves 16 bytes per load (a 1 KiB row per wave instruction);
// otherwise 8.  Result rows are written with non-temporal stores when nt is set
// (they are not read again here).
template <int U>
__global__ void __launch_bounds__(256) contains_kernel(IbfView g,
                                                       uint64_t const * __restrict__ values,
                                                       uint64_t n,
                                                       uint64_t * __restrict__ out,
                                                       uint32_t vec,
                                                       uint32_t nt)
{
    typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
    int const lane = static_cast<int>(threadIdx.x & 63u);
    uint64_t const waves = static_cast<uint64_t>(gridDim.x) * (blockDim.x >> 6);
    uint64_t const bw = g.bin_words;
    uint32_t const hf = g.hash_funs;
    for (uint64_t base = (static_cast<uint64_t>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64u;
         base < n; base += waves * 64u)
    {
        uint32_t const nv = static_cast<uint32_t>(min<uint64_t>(64u, n - base));
        uint64_t const v = static_cast<uint32_t>(lane) < nv ? values[base + lane] : 0ULL;
        uint64_t rows[5];
#pragma unroll
        for (int k = 0; k < 5; ++k)
            rows[k] = static_cast<uint32_t>(k) < hf ? row_word(v, hash_seed(k), g.hash_shift, g.bin_size, bw) : 0ULL;
        for (uint32_t i = 0; i < nv; i += U)
        {
            uint64_t r[U][5];
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int k = 0; k < 5; ++k)
                    r[u][k] = static_cast<uint32_t>(k) < hf ? readlane64(rows[k], static_cast<int>(min(i + u, 63u)))
                                                            : 0ULL;
            if (vec)
            {
                for (uint64_t w = 2u * static_cast<uint64_t>(lane); w < bw; w += 128u)
                {
                    u64x2 acc[U];
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        acc[u] = *reinterpret_cast<u64x2 const *>(g.words + r[u][0] + w);
#pragma unroll
                    for (int k = 1; k < 5; ++k)
                        if (static_cast<uint32_t>(k) < hf)
#pragma unroll
                            for (int u = 0; u < U; ++u)
                                acc[u] &= *reinterpret_cast<u64x2 const *>(g.words + r[u][k] + w);
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        if (i + u < nv)
                        {
                            u64x2 * const o = reinterpret_cast<u64x2 *>(out + (base + i + u) * bw + w);
                            if (nt)
                                __builtin_nontemporal_store(acc[u], o);
                            else
                                *o = acc[u];
                        }
                }
            }
            else
            {
                for (uint64_t w = static_cast<uint64_t>(lane); w < bw; w += 64u)
                {
                    uint64_t acc[U];
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        acc[u] = g.words[r[u][0] + w];
#pragma unroll
                    for (int k = 1; k < 5; ++k)
                        if (static_cast<uint32_t>(k) < hf)
#pragma unroll
                            for (int u = 0; u < U; ++u)
                                acc[u] &= g.words[r[u][k] + w];
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        if (i + u < nv)
                            out[(base + i + u) * bw + w] = acc[u];
                }
            }
        }
    }
}

template <typename T>
__device__ __forceinline__ void store_counts(T * __restrict__ out, uint32_t const * lds, uint32_t n, int tid, int nthreads)
{
    for (uint32_t i = static_cast<uint32_t>(tid); i < n; i += static_cast<uint32_t>(nthreads))
        out[i] = static_cast<T>(lds[i]);
}

// Batched counting_agent_type<T>::bulk_count (interleaved_bloom_filter.hpp:652-665).
// One block per query (grid-stride); the block's waves split the query's
// values, accumulate per-bin counters of one word tile in LDS, then write the
// tile's counters truncated to counter_bytes.
template <int LPV, int CW>
__global__ void __launch_bounds__(256) count_kernel(IbfView g,
                                                    uint64_t const * __restrict__ values,
                                                    uint64_t const * __restrict__ offsets,
                                                    uint64_t nq,
                                                    uint32_t tile_words,
                                                    uint32_t counter_bytes,
                                                    void * __restrict__ counts)
{
    extern __shared__ uint32_t lds[];
    int const tid = static_cast<int>(threadIdx.x);
    int const nthreads = static_cast<int>(blockDim.x);
    int const lane = tid & 63;
    int const wave = tid >> 6;
    int const nwaves = nthreads >> 6;

    for (uint64_t q = blockIdx.x; q < nq; q += gridDim.x)
    {
        uint64_t const beg = offsets[q];
        uint32_t const L = static_cast<uint32_t>(offsets[q + 1] - beg);
        for (uint32_t w0 = 0; w0 < g.bin_words; w0 += tile_words)
        {
            uint32_t const tw = min(tile_words, g.bin_words - w0);
            for (uint32_t i = static_cast<uint32_t>(tid); i < tw * 64u; i += static_cast<uint32_t>(nthreads))
                lds[i] = 0u;
            __syncthreads();
            count_into_lds<LPV, CW, 2>(g, values + beg, L, w0, tw, lds, wave, nwaves, lane);
            __syncthreads();
            uint32_t const first_bin = w0 * 64u;
            uint32_t const nb = min(tw * 64u, g.bins - first_bin);
            uint64_t const o = q * g.bins + first_bin;
            switch (counter_bytes)
            {
            case 1: store_counts(static_cast<uint8_t *>(counts) + o, lds, nb, tid, nthreads); break;
            case 2: store_counts(static_cast<uint16_t *>(counts) + o, lds, nb, tid, nthreads); break;
            case 4: store_counts(static_cast<uint32_t *>(counts) + o, lds, nb, tid, nthreads); break;
            default: store_counts(static_cast<uint64_t *>(counts) + o, lds, nb, tid, nthreads); break;
            }
            __syncthreads();
        }
    }
}

// increase_bin_number_to (src/interleaved_bloom_filter.cpp:163-201): row r of
// the old matrix (old_bw words) becomes row r of the new one (new_bw words);
// the added words of every row are zero.  One thread per new word.
__global__ void __launch_bounds__(256) restride_kernel(uint64_t const * __restrict__ old_words,
                                                       uint64_t old_bw,
                                                       uint64_t * __restrict__ new_words,
                                                       uint64_t new_bw,
                                                       uint64_t n_new)
{
    uint64_t const stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n_new; i += stride)
    {
        uint64_t const r = i / new_bw;
        uint64_t const w = i - r * new_bw;
        new_words[i] = w < old_bw ? old_words[r * old_bw + w] : 0ULL;
    }
}

namespace
{
inline unsigned grid_for(uint64_t items, unsigned per_block, unsigned cap = 1u << 20)
{
    uint64_t b = (items + per_block - 1) / per_block;
    if (b == 0)
        b = 1;
    return static_cast<unsigned>(b < cap ? b : cap);
}
} // namespace

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
                          hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(emplace_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, words, bin_size, bin_words,
                       hash_shift, hash_funs, values, bins, bin, n, bin_count);
    return hipGetLastError();
}

hipError_t launch_emplace_positions(uint64_t bin_size, uint32_t bin_words, uint32_t hash_shift, uint32_t hash_funs,
                                    uint64_t const * values, uint64_t const * bins, uint64_t bin, uint64_t n,
                                    uint64_t bin_count, uint64_t invalid_key, uint64_t * keys, uint64_t * owners,
                                    hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(emplace_positions_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, bin_size, bin_words,
                       hash_shift, hash_funs, values, bins, bin, n, bin_count, invalid_key, keys, owners);
    return hipGetLastError();
}

hipError_t launch_mark_first_touches(uint64_t const * words, uint64_t const * keys, uint64_t const * owners,
                                     uint64_t m, uint64_t invalid_key, uint8_t * marked, hipStream_t stream)
{
    if (m == 0)
        return hipSuccess;
    hipLaunchKernelGGL(mark_first_touches_kernel, dim3(grid_for(m, 256)), dim3(256), 0, stream, words, keys, owners,
                       m, invalid_key, marked);
    return hipGetLastError();
}

hipError_t launch_count_marked(uint8_t const * marked, uint64_t const * bins, uint64_t bin, uint64_t n,
                               unsigned long long * occupancy, hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(count_marked_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, marked, bins, bin, n,
                       occupancy);
    return hipGetLastError();
}

hipError_t launch_emplace_synth(uint64_t * words,
                                uint64_t bin_size,
                                uint32_t bin_words,
                                uint32_t hash_shift,
                                uint32_t hash_funs,
                                uint64_t seed,
                                uint64_t first_bin,
                                uint64_t n_bins,
                                uint64_t per_bin,
                                hipStream_t stream)
{
    uint64_t const n = n_bins * per_bin;
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(emplace_synth_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, words, bin_size,
                       bin_words, hash_shift, hash_funs, seed, first_bin, per_bin, n);
    return hipGetLastError();
}

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
                                  hipStream_t stream)
{
    if (n_chunks == 0)
        return hipSuccess;
    unsigned const grid = static_cast<unsigned>(n_chunks < (1u << 20) ? n_chunks : (1u << 20));
    hipLaunchKernelGGL(segment_emplace_kernel, dim3(grid), dim3(256), 0, stream, desc, arena, segs, chunks, n_chunks,
                       values, ub_offsets, seed, chunk_ids, chunk_id_bytes);
    return hipGetLastError();
}

hipError_t launch_restride(uint64_t const * old_words,
                           uint64_t old_bw,
                           uint64_t * new_words,
                           uint64_t new_bw,
                           uint64_t bin_size,
                           hipStream_t stream)
{
    uint64_t const n = new_bw * bin_size;
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(restride_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, old_words, old_bw, new_words,
                       new_bw, n);
    return hipGetLastError();
}

hipError_t launch_clear(uint64_t * words,
                        uint64_t bin_size,
                        uint32_t bin_words,
                        uint64_t const * d_bins,
                        uint32_t nbins,
                        hipStream_t stream)
{
    if (nbins == 0)
        return hipSuccess;
    hipLaunchKernelGGL(clear_kernel, dim3(grid_for(bin_size, 256)), dim3(256), 0, stream, words, bin_size,
                       bin_words, d_bins, nbins);
    return hipGetLastError();
}

hipError_t launch_contains(IbfView const & g, uint64_t const * values, uint64_t n, uint64_t * out, hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    uint32_t const vec = g.bin_words % 2u == 0u && reinterpret_cast<uintptr_t>(g.words) % 16u == 0u &&
                         reinterpret_cast<uintptr_t>(out) % 16u == 0u;
    uint32_t const nt = g_count_tuning.nt_stores;
    unsigned const grid = grid_for(n, 256);
    switch (g_count_tuning.contains_unroll)
    {
    case 1: hipLaunchKernelGGL(contains_kernel<1>, dim3(grid), dim3(256), 0, stream, g, values, n, out, vec, nt); break;
    case 2: hipLaunchKernelGGL(contains_kernel<2>, dim3(grid), dim3(256), 0, stream, g, values, n, out, vec, nt); break;
    case 8: hipLaunchKernelGGL(contains_kernel<8>, dim3(grid), dim3(256), 0, stream, g, values, n, out, vec, nt); break;
    default: hipLaunchKernelGGL(contains_kernel<4>, dim3(grid), dim3(256), 0, stream, g, values, n, out, vec, nt);
    }
    return hipGetLastError();
}

// Shape of the count kernel for an IBF: tiles of at most 128 words (8192
// technical bins, 32 KiB of LDS counters).
void count_shape(uint32_t bin_words, uint32_t & lpv_log, uint32_t & cw, uint32_t & tile_words)
{
    if (bin_words > 64u)
    {
        lpv_log = 6u;
        cw = 2u;
        tile_words = 128u;
        return;
    }
    uint32_t l = 0;
    while ((1u << l) < bin_words)
        ++l;
    lpv_log = l;
    cw = 1u;
    tile_words = bin_words;
}

hipError_t launch_count(IbfView const & g,
                        uint64_t const * values,
                        uint64_t const * offsets,
                        uint64_t nq,
                        uint32_t counter_bytes,
                        void * counts,
                        unsigned max_blocks,
                        hipStream_t stream)
{
    if (nq == 0)
        return hipSuccess;
    if (count_wide_applies(g))
        return launch_count_wide(g, values, offsets, nq, counter_bytes, counts, max_blocks, stream);
    uint32_t lpv_log, cw, tile_words;
    count_shape(g.bin_words, lpv_log, cw, tile_words);
    size_t const lds = static_cast<size_t>(tile_words) * 64u * sizeof(uint32_t);
    unsigned const grid = static_cast<unsigned>(nq < max_blocks ? nq : max_blocks);
    dim3 const blk(256);
#define HIBF_COUNT_CASE(LL, LPVV, CWW)                                                                                 \
    case LL:                                                                                                           \
        hipLaunchKernelGGL((count_kernel<LPVV, CWW>), dim3(grid), blk, lds, stream, g, values, offsets, nq,           \
                           tile_words, counter_bytes, counts);                                                        \
        break;
    if (cw == 2u)
    {
        hipLaunchKernelGGL((count_kernel<64, 2>), dim3(grid), blk, lds, stream, g, values, offsets, nq, tile_words,
                           counter_bytes, counts);
    }
    else
    {
        switch (lpv_log)
        {
            HIBF_COUNT_CASE(0, 1, 1)
            HIBF_COUNT_CASE(1, 2, 1)
            HIBF_COUNT_CASE(2, 4, 1)
            HIBF_COUNT_CASE(3, 8, 1)
            HIBF_COUNT_CASE(4, 16, 1)
            HIBF_COUNT_CASE(5, 32, 1)
        default:
            hipLaunchKernelGGL((count_kernel<64, 1>), dim3(grid), blk, lds, stream, g, values, offsets, nq,
                               tile_words, counter_bytes, counts);
            break;
        }
    }
#undef HIBF_COUNT_CASE
    return hipGetLastError();
}

} // namespace hibf_amd
