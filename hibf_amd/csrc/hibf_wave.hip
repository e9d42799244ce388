// HIBF descent, one wavefront per query (gfx950).
//
// Each wave of a block is independent: it steals a query from the agent's
// cursor, walks the HIBF depth first with per-wave LDS state (counters of the
// IBF being counted, per-level event lists, a result buffer), and never waits
// for the other waves of its block.  Counting a narrow IBF (<= 4 words per
// row, e.g. 256 technical bins) gives each lane a whole value: the lane
// loads the h rows of its value (up to 4 words each) and adds the set bits
// of their AND into the LDS counters, so a wave keeps 2 x 64 values' rows in
// flight per round trip.  Wider IBFs use the lanes-per-value scheme of the
// flat kernels.  Results follow the reference's recursion order
// (hierarchical_interleaved_bloom_filter.hpp:321-355, :462-496).
//
// Membership ids of a query are buffered in LDS (kWaveResults entries); a
// query with more hits is recorded in an overflow list and answered again by
// the block-per-query kernel, which keeps a full-size list per block.
#include "descent_common.hpp"
#include "device_common.hpp"
#include "kernels.hpp"

namespace hibf_amd
{

namespace
{

constexpr uint32_t kWaveResults = 128;

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Lane-per-value counting for IBFs of BW <= 4 words per row: B batches of
// 64 values in flight per round.
template <int BW, int B>
__device__ __forceinline__ void count_narrow(IbfView const & g,
                                             uint64_t const * __restrict__ vals,
                                             uint32_t L,
                                             uint32_t * cnt,
                                             int lane)
{
    uint32_t const hf = g.hash_funs;
    uint64_t const last_mask = tail_mask(g, static_cast<uint32_t>(BW - 1));
    for (uint32_t base = 0; base < L; base += 64u * B)
    {
        uint64_t x[B][BW];
#pragma unroll
        for (int b = 0; b < B; ++b)
        {
            uint32_t const i = base + static_cast<uint32_t>(b * 64 + lane);
            bool const valid = i < L;
            uint64_t const v = valid ? vals[i] : 0ULL;
#pragma unroll
            for (int w = 0; w < BW; ++w)
                x[b][w] = 0ULL;
            if (valid)
            {
                uint64_t const r0 = row_word(v, hash_seed(0), g.hash_shift, g.bin_size, BW);
#pragma unroll
                for (int w = 0; w < BW; ++w)
                    x[b][w] = g.words[r0 + w];
#pragma unroll
                for (int k = 1; k < 5; ++k)
                    if (static_cast<uint32_t>(k) < hf)
                    {
                        uint64_t const rk = row_word(v, hash_seed(k), g.hash_shift, g.bin_size, BW);
#pragma unroll
                        for (int w = 0; w < BW; ++w)
                            x[b][w] &= g.words[rk + w];
                    }
                x[b][BW - 1] &= last_mask;
            }
        }
#pragma unroll
        for (int b = 0; b < B; ++b)
        {
#pragma unroll
            for (int w = 0; w < BW; ++w)
            {
                uint64_t bits = x[b][w];
                while (bits)
                {
                    atomicAdd(&cnt[w * 64 + __builtin_ctzll(bits)], 1u);
                    bits &= bits - 1ULL;
                }
            }
        }
    }
}

// HF == 2 (every IBF with two hash functions, rows of at most 4 words):
// lanes share a value's row (count_into_lds_dispatch_narrow), so a row is one
// load instruction's worth of lanes and no line is fetched twice.
template <bool NARROW, int HF = 0>
__device__ __forceinline__ void count_visit(HibfIbfDesc const & d,
                                            uint64_t const * arena,
                                            uint64_t const * __restrict__ vals,
                                            uint32_t L,
                                            uint32_t * cnt,
                                            int lane)
{
    IbfView const g{arena + d.word_offset, d.bin_size, d.bins, d.bin_words, d.hash_shift, d.hash_funs};
    if constexpr (NARROW && HF == 2)
    {
        count_into_lds_dispatch_narrow<2, false, 2>(d.lpv_log, g, vals, L, cnt, 0, 1, lane);
        return;
    }
    switch (d.bin_words)
    {
    case 1: count_narrow<1, 2>(g, vals, L, cnt, lane); break;
    case 2: count_narrow<2, 2>(g, vals, L, cnt, lane); break;
    case 3: count_narrow<3, 2>(g, vals, L, cnt, lane); break;
    case 4: count_narrow<4, 2>(g, vals, L, cnt, lane); break;
    default:
        if constexpr (!NARROW)
            count_into_lds_dispatch<2>(d.lpv_log, g, vals, L, cnt, 0, 1, lane);
        break;
    }
}


// NARROW: every IBF of the HIBF has at most 4 words per row (checked on the host).
template <bool COUNTING, bool NARROW, int HF = 0>
__device__ __forceinline__ void wave_body(HibfLaunch const & h,
                                                        uint64_t const * __restrict__ values,
                                                        uint64_t const * __restrict__ offsets,
                                                        uint64_t nq,
                                                        uint16_t const * __restrict__ thr16,
                                                        uint64_t const * __restrict__ thr64,
                                                        uint32_t counter_bytes,
                                                        void * __restrict__ counts,
                                                        uint64_t * __restrict__ ids,
                                                        uint64_t ids_cap,
                                                        uint64_t * __restrict__ hit_start,
                                                        uint64_t * __restrict__ hit_count,
                                                        uint64_t * __restrict__ overflow)
{
    extern __shared__ uint32_t smem[];
    int const lane = static_cast<int>(threadIdx.x & 63u);
    int const wave = static_cast<int>(threadIdx.x >> 6);
    // Per-wave LDS: counters, event lists, results, level state.
    uint32_t const per_wave = h.max_bins * (1u + h.depth) + kWaveResults + 3u * kHibfMaxDepth;
    uint32_t * const cnt = smem + static_cast<uint32_t>(wave) * per_wave;
    uint32_t * const evs = cnt + h.max_bins;
    uint32_t * const res = evs + h.max_bins * h.depth;
    uint32_t * const nev = res + kWaveResults;
    uint32_t * const pos = nev + kHibfMaxDepth;
    uint32_t * const ibf_at = pos + kHibfMaxDepth;

    uint64_t visits = 0, lookups = 0, row_bytes = 0;
    while (true)
    {
        uint64_t q = 0;
        if (lane == 0)
            q = atomicAdd(&h.counters[1], 1ULL);
        q = readlane64(q, 0);
        if (q >= nq)
            break;
        uint64_t const beg = offsets[q];
        uint32_t const L = static_cast<uint32_t>(offsets[q + 1] - beg);
        uint64_t const * const vals = values + beg;
        uint64_t const threshold = COUNTING ? thr64[q] : static_cast<uint64_t>(thr16[q]);
        uint32_t const cb = COUNTING ? counter_bytes : 2u;
        void * const row = COUNTING ? static_cast<void *>(static_cast<char *>(counts) + q * h.n_user_bins * cb)
                                    : nullptr;
        if constexpr (COUNTING)
        {
            zero_counter_row(row, h.n_user_bins, cb, h.nt_stores != 0u, lane, 64);
        }
        uint32_t hits = 0;
        bool spilled = false;
        uint32_t depth = 0;
        uint32_t ibf = 0;
        bool descend = true;
        while (true)
        {
            if (descend)
            {
                HibfIbfDesc const d = h.ibfs[ibf];
                for (uint32_t i = static_cast<uint32_t>(lane); i < d.bins; i += 64u)
                    cnt[i] = 0u;
                wave_sync();
                count_visit<NARROW, HF>(d, h.arena, vals, L, cnt, lane);
                ++visits;
                lookups += L;
                row_bytes += static_cast<uint64_t>(L) * d.hash_funs * d.bin_words * 8u;
                wave_sync();
                uint32_t const n =
                    scan_ibf<COUNTING>(d, h.tb_info, cnt, threshold, cb, evs + depth * h.max_bins, row, lane);
                if (lane == 0)
                {
                    nev[depth] = n;
                    pos[depth] = 0;
                    ibf_at[depth] = ibf;
                }
                wave_sync();
                descend = false;
            }
            uint32_t const p = pos[depth];
            if (p < nev[depth])
            {
                uint32_t const e = evs[depth * h.max_bins + p];
                bool const merged = (e & kEvMerged) != 0u;
                uint32_t const payload = e & ~kEvMerged;
                if (lane == 0)
                {
                    pos[depth] = p + 1u;
                    if (!COUNTING && !merged)
                    {
                        if (hits < kWaveResults)
                            res[hits] = payload;
                        else
                            spilled = true;
                    }
                }
                if (!COUNTING && !merged)
                    ++hits;
                wave_sync();
                if (merged)
                {
                    ++depth;
                    ibf = payload;
                    descend = true;
                }
                continue;
            }
            if (depth == 0)
                break;
            --depth;
            ibf = ibf_at[depth];
        }
        if constexpr (!COUNTING)
        {
            // With a per-query cap that the result buffer holds, the first cap ids are all that is kept.
            uint32_t const cap = h.per_query_cap;
            spilled = hits > kWaveResults && !(cap != 0u && cap <= kWaveResults);
            uint32_t const kept = cap ? min(hits, cap) : hits;
            uint64_t start = 0;
            if (lane == 0)
            {
                if (spilled)
                {
                    // answered again by the block kernel; hit_count is set there
                    uint64_t const k = atomicAdd(&h.counters[6], 1ULL);
                    overflow[k] = q;
                    hit_start[q] = 0;
                    hit_count[q] = 0;
                }
                else
                {
                    uint64_t const found = atomicAdd(&h.counters[0], static_cast<unsigned long long>(hits));
                    start = cap ? q * cap : found;
                    hit_start[q] = start;
                    hit_count[q] = hits;
                    if (start + kept > ids_cap)
                        h.counters[5] = 1ULL;
                }
            }
            start = readlane64(start, 0);
            if (!spilled)
                for (uint32_t i = static_cast<uint32_t>(lane); i < kept; i += 64u)
                    if (start + i < ids_cap)
                        ids[start + i] = res[i];
            wave_sync();
        }
    }
    if (lane == 0 && visits)
    {
        atomicAdd(&h.counters[2], static_cast<unsigned long long>(visits));
        atomicAdd(&h.counters[3], static_cast<unsigned long long>(row_bytes));
        atomicAdd(&h.counters[4], static_cast<unsigned long long>(lookups));
    }
}

#define HIBF_WAVE_PARAMS                                                                                             \
    HibfLaunch h, uint64_t const * __restrict__ values, uint64_t const * __restrict__ offsets, uint64_t nq,            \
        uint16_t const * __restrict__ thr16, uint64_t const * __restrict__ thr64, uint32_t counter_bytes,               \
        void * __restrict__ counts, uint64_t * __restrict__ ids, uint64_t ids_cap, uint64_t * __restrict__ hit_start,  \
        uint64_t * __restrict__ hit_count, uint64_t * __restrict__ overflow
#define HIBF_WAVE_ARGS                                                                                               \
    h, values, offsets, nq, thr16, thr64, counter_bytes, counts, ids, ids_cap, hit_start, hit_count, overflow

template <bool COUNTING, bool NARROW>
__global__ void __launch_bounds__(256) hibf_wave_kernel(HIBF_WAVE_PARAMS)
{
    wave_body<COUNTING, NARROW>(HIBF_WAVE_ARGS);
}

// Membership with two hash functions over rows of at most 4 words.
__global__ void __launch_bounds__(256) hibf_wave_membership_h2_w4_kernel(HIBF_WAVE_PARAMS)
{
    wave_body<false, true, 2>(HIBF_WAVE_ARGS);
}

#undef HIBF_WAVE_PARAMS
#undef HIBF_WAVE_ARGS

} // namespace

size_t hibf_wave_lds_bytes(HibfLaunch const & h, unsigned waves_per_block)
{
    return static_cast<size_t>(waves_per_block) *
           (h.max_bins * (1u + h.depth) + kWaveResults + 3u * kHibfMaxDepth) * sizeof(uint32_t);
}

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
                                       hipStream_t stream)
{
    if (nq == 0)
        return hipSuccess;
    if (h.max_bin_words <= 4u && h.uniform_hash_funs == 2u)
        hipLaunchKernelGGL(hibf_wave_membership_h2_w4_kernel, dim3(blocks), dim3(256), hibf_wave_lds_bytes(h, 4),
                           stream, h, values, offsets, nq, thresholds, nullptr, 2u, nullptr, ids, ids_cap, hit_start,
                           hit_count, overflow);
    else if (h.max_bin_words <= 4u)
        hipLaunchKernelGGL((hibf_wave_kernel<false, true>), dim3(blocks), dim3(256), hibf_wave_lds_bytes(h, 4), stream,
                           h, values, offsets, nq, thresholds, nullptr, 2u, nullptr, ids, ids_cap, hit_start,
                           hit_count, overflow);
    else
        hipLaunchKernelGGL((hibf_wave_kernel<false, false>), dim3(blocks), dim3(256), hibf_wave_lds_bytes(h, 4),
                           stream, h, values, offsets, nq, thresholds, nullptr, 2u, nullptr, ids, ids_cap, hit_start,
                           hit_count, overflow);
    return hipGetLastError();
}

hipError_t launch_hibf_wave_count(HibfLaunch const & h,
                                  uint64_t const * values,
                                  uint64_t const * offsets,
                                  uint64_t nq,
                                  uint64_t const * thresholds,
                                  uint32_t counter_bytes,
                                  void * counts,
                                  unsigned blocks,
                                  hipStream_t stream)
{
    if (nq == 0)
        return hipSuccess;
    if (h.max_bin_words <= 4u)
        hipLaunchKernelGGL((hibf_wave_kernel<true, true>), dim3(blocks), dim3(256), hibf_wave_lds_bytes(h, 4), stream,
                           h, values, offsets, nq, nullptr, thresholds, counter_bytes, counts, nullptr, 0ULL, nullptr,
                           nullptr, nullptr);
    else
        hipLaunchKernelGGL((hibf_wave_kernel<true, false>), dim3(blocks), dim3(256), hibf_wave_lds_bytes(h, 4),
                           stream, h, values, offsets, nq, nullptr, thresholds, counter_bytes, counts, nullptr, 0ULL,
                           nullptr, nullptr, nullptr);
    return hipGetLastError();
}

} // namespace hibf_amd
