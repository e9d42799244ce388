// Helpers shared by the two HIBF descent kernels (query per block in
// hibf_kernels.hip, query per wave in hibf_wave.hip): counter stores of the
// counting agent's output rows, its zero fill, and the technical-bin scan that
// turns one IBF's counts into events.
#pragma once

#include "device_common.hpp"
#include "kernels.hpp"

namespace hibf_amd
{

__device__ __forceinline__ uint64_t width_mask(uint32_t counter_bytes)
{
    return counter_bytes >= 8u ? ~0ULL : ((1ULL << (8u * counter_bytes)) - 1ULL);
}

__device__ __forceinline__ void store_counter(void * base, uint64_t i, uint32_t counter_bytes, uint64_t v)
{
    switch (counter_bytes)
    {
    case 1: static_cast<uint8_t *>(base)[i] = static_cast<uint8_t>(v); break;
    case 2: static_cast<uint16_t *>(base)[i] = static_cast<uint16_t>(v); break;
    case 4: static_cast<uint32_t *>(base)[i] = static_cast<uint32_t>(v); break;
    default: static_cast<uint64_t *>(base)[i] = v; break;
    }
}

// counting_vector fill(0) of one output row (hierarchical_interleaved_bloom_filter.hpp:561)
// by threads tid, tid + nthreads, ...: 16-byte stores (non-temporal when nt) over
// the 16-byte-aligned middle of the row, counter stores at its ends.  The row
// starts at a multiple of cb and cb divides 16, so both ends are whole counters.
__device__ __forceinline__ void zero_counter_row(void * row, uint64_t n, uint32_t cb, bool nt, int tid, int nthreads)
{
    char * const bytes = static_cast<char *>(row);
    uint64_t const len = n * cb;
    uint64_t const head = (16u - (reinterpret_cast<uintptr_t>(bytes) & 15u)) & 15u;
    uint64_t const n16 = len > head ? (len - head) / 16u : 0u;
    uint64_t const tail = head + n16 * 16u;
    if (nt)
    {
        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
        u32x4 * const mid = reinterpret_cast<u32x4 *>(bytes + head);
        for (uint64_t i = static_cast<uint64_t>(tid); i < n16; i += static_cast<uint64_t>(nthreads))
            __builtin_nontemporal_store(u32x4{0u, 0u, 0u, 0u}, mid + i);
    }
    else
    {
        uint4 * const mid = reinterpret_cast<uint4 *>(bytes + head);
        for (uint64_t i = static_cast<uint64_t>(tid); i < n16; i += static_cast<uint64_t>(nthreads))
            mid[i] = make_uint4(0u, 0u, 0u, 0u);
    }
    uint64_t const n_head = (head < len ? head : len) / cb;
    for (uint64_t i = static_cast<uint64_t>(tid); i < n_head; i += static_cast<uint64_t>(nthreads))
        store_counter(row, i, cb, 0ULL);
    uint64_t const first_tail = tail / cb > n_head ? tail / cb : n_head;
    for (uint64_t i = first_tail + static_cast<uint64_t>(tid); i < n; i += static_cast<uint64_t>(nthreads))
        store_counter(row, i, cb, 0ULL);
}

// Event entries of scan_ibf: the payload of the technical bin (user-bin id,
// or child IBF of a merged bin) with kEvMerged set for merged bins, so that
// processing an event reads nothing but LDS.  Device ids are below 2^31.
constexpr uint32_t kEvMerged = 0x80000000u;

// Scan the technical bins of the IBF just counted into `cnt` (one whole wave),
// exactly as the loop of hierarchical_interleaved_bloom_filter.hpp:329-354:
// the running sum adds every bin's count (deleted bins included), resets at a
// merged bin and at the end of a user bin's run, and a reset point is an event
// iff the sum (modulo 2^(8*counter_bytes)) reaches the threshold.  Events are
// appended to ev in technical-bin order.  In counting mode user-bin events are
// written straight to the output row and only merged bins become events.
// The table entries of up to four 64-bin chunks are loaded before the first
// of them is scanned, so a 256-bin IBF costs one memory latency.
// Returns the number of events.
template <bool COUNTING>
__device__ __forceinline__ uint32_t scan_ibf(HibfIbfDesc const & d,
                                             uint64_t const * __restrict__ tb_info,
                                             uint32_t const * cnt,
                                             uint64_t threshold,
                                             uint32_t counter_bytes,
                                             uint32_t * ev,
                                             void * count_row,
                                             int lane)
{
    uint64_t const mask = width_mask(counter_bytes);
    uint64_t carry = 0;
    uint32_t nev = 0;
    for (uint32_t g0 = 0; g0 < d.bins; g0 += 256u)
    {
        uint64_t infos[4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            uint32_t const tb = g0 + 64u * static_cast<uint32_t>(j) + static_cast<uint32_t>(lane);
            infos[j] = tb < d.bins ? tb_info[d.tb_offset + tb] : 0ULL;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            uint32_t const c0 = g0 + 64u * static_cast<uint32_t>(j);
            if (c0 >= d.bins)
                break;
            uint32_t const tb = c0 + static_cast<uint32_t>(lane);
            bool const valid = tb < d.bins;
            uint64_t const info = infos[j];
            uint32_t const kind = static_cast<uint32_t>(info >> 32);
            uint32_t const payload = static_cast<uint32_t>(info);
            bool const reset = valid && (kind == kTbRunEnd || kind == kTbMerged);
            uint64_t sum;
            if (counter_bytes <= 4u)
            {
                // Sums modulo 2^32 suffice for counters of up to 4 bytes: a run's sum is the difference of two
                // plain prefix sums (the one at its reset point and the one at the reset point before it), or
                // the carried sum plus the prefix when no reset precedes it in this chunk.
                uint32_t const p = wave_inclusive_add_u32(valid ? cnt[tb] : 0u);
                uint64_t const resets = __ballot(reset);
                uint64_t const before = resets & ((1ULL << lane) - 1ULL);
                int const prev = before ? 63 - __builtin_clzll(before) : lane;
                uint32_t const at_prev = static_cast<uint32_t>(__shfl(static_cast<int>(p), prev, 64));
                sum = static_cast<uint64_t>(before ? p - at_prev : static_cast<uint32_t>(carry) + p) & mask;
                uint32_t const total = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(p), 63));
                if (resets)
                {
                    int const last = 63 - __builtin_clzll(resets);
                    carry = total - static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(p), last));
                }
                else
                {
                    carry = static_cast<uint32_t>(carry) + total;
                }
            }
            else
            {
                uint64_t s = valid ? (static_cast<uint64_t>(cnt[tb]) & mask) : 0ULL;
                // Segment heads: the bin after a reset point.
                bool const prev_reset = __shfl_up(static_cast<int>(reset), 1, 64) != 0;
                bool f = lane > 0 && prev_reset;
#pragma unroll
                for (int dlt = 1; dlt < 64; dlt <<= 1)
                {
                    uint64_t const y = shfl64(s, lane >= dlt ? lane - dlt : lane);
                    bool const fy = __shfl(static_cast<int>(f), lane >= dlt ? lane - dlt : lane, 64) != 0;
                    if (lane >= dlt)
                    {
                        if (!f)
                            s += y;
                        f = f || fy;
                    }
                }
                if (!f)
                    s += carry;
                sum = s & mask;
                // Pending sum after this chunk: lane 63's segment sum unless it is a reset point.
                uint64_t const last_s = shfl64(s, 63);
                bool const last_reset = __shfl(static_cast<int>(reset), 63, 64) != 0;
                carry = last_reset ? 0ULL : last_s;
            }
            bool const hit = reset && sum >= threshold;
            bool is_event = hit;
            if constexpr (COUNTING)
            {
                // A user bin whose run is cut (by a deleted bin) keeps the sum of its last run that reached the
                // threshold (hierarchical_interleaved_bloom_filter.hpp:488-491): of two such run ends in this
                // chunk, the higher bin stores.
                if (highest_lane_per_id(hit && kind == kTbRunEnd, payload, lane))
                    store_counter(count_row, payload, counter_bytes, sum);
                is_event = hit && kind == kTbMerged;
            }
            uint64_t const ballot = __ballot(is_event);
            uint32_t const before = static_cast<uint32_t>(__popcll(ballot & ((1ULL << lane) - 1ULL)));
            if (is_event)
                ev[nev + before] = payload | (kind == kTbMerged ? kEvMerged : 0u);
            nev += static_cast<uint32_t>(__popcll(ballot));
        }
    }
    return nev;
}

} // namespace hibf_amd
